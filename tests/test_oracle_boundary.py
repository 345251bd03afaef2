"""Pins the oracle's boundary-particle terms (RX_USE_BOUNDARY_PARTICLE: rx_sph_cpu.cpp:527-693,
rx_sph_pbd.cpp:699-1026, rx_sph_kernel.cu:662-727) without retyping them.

With V_b = m / rho0 a boundary particle acts on the fluid exactly as a fluid particle of mass m would:
the density, pressure-force (with p_b = 0), multiplier (where the boundary gradients cancel) and
position-correction terms must then equal what the oracle's fluid-only functions give for the union of the
two sets.  The volume is pinned by a planar lattice sum, the boundary multiplier by a hand-worked pair."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import ptr
from test_oracle_pins import make_params, pos4, poly6


def spiky_grad_mag(r, h):  # |grad W_spiky| (Mueller et al. 2003, eq. (21))
    return 45.0 / (math.pi * h ** 6) * (h - r) ** 2


def _cloud(n, seed, lo=0.3, hi=0.7):
    rng = np.random.default_rng(seed)
    return pos4(rng.uniform(lo, hi, (n, 3)))


def _wall(h, z=0.25, n=14, s=None):
    """A planar square lattice of boundary particles at height z (spacing h/2)."""
    s = s or 0.5 * h
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 2) * s + 0.5 - 0.5 * (n - 1) * s
    return pos4(np.column_stack([g, np.full(len(g), z)]))


def test_volume_of_planar_lattice(orc):
    # interior particle of a plane lattice at spacing s: V = 1 / sum_{a,b} W(s sqrt(a^2 + b^2)), independent of m
    h, s = 0.1, 0.04
    n = 21
    b = _wall(h, n=n, s=s)
    vol = np.zeros(len(b), np.float32)
    for mass in (0.125, 2.0):
        P = make_params(h=h, mass=mass)
        orc.orc_boundary_volume_f(ptr(b), len(b), C.byref(P), ptr(vol))
        k = np.arange(-5, 6)
        a, c = np.meshgrid(k, k, indexing="ij")
        expect = 1.0 / poly6(s * np.sqrt(a * a + c * c), h).sum()
        centre = (n // 2) * n + n // 2
        assert vol[centre] == pytest.approx(expect, rel=2e-6)
        # an edge particle has fewer neighbours, hence a larger volume
        assert vol[n // 2] > 1.2 * vol[centre]


def _union_params(h=0.1):
    return make_params(h=h, mass=0.125, rest=998.29, gas_k=3.0, lo=0.0, hi=1.0, use_ap=1, eps=600.0)


def test_density_term_equals_fluid_particles_of_mass_m(orc):
    P = _union_params()
    f = _cloud(300, 1)
    b = _wall(P.h)
    vol = np.full(len(b), P.mass / P.rest_dens, np.float32)
    dens = np.zeros(len(f), np.float32)
    orc.orc_pbf_density_f(ptr(f), len(f), C.byref(P), ptr(dens))
    base = dens.copy()
    pres = np.zeros(len(f), np.float32)
    orc.orc_boundary_density_f(ptr(f), len(f), ptr(b), ptr(vol), len(b), C.byref(P), ptr(dens), ptr(pres))
    u = np.concatenate([f, b])
    du = np.zeros(len(u), np.float32)
    orc.orc_pbf_density_f(ptr(u), len(u), C.byref(P), ptr(du))
    touched = dens != base
    assert touched.sum() > 20  # the wall is within h of part of the cloud
    np.testing.assert_allclose(dens, du[:len(f)], rtol=2e-6)
    np.testing.assert_allclose(pres, np.float32(P.gas_k) * (dens - np.float32(P.rest_dens)), rtol=1e-5, atol=1e-3)


def test_pressure_term_equals_fluid_particles_at_rest_density(orc):
    # boundary-as-fluid with rho_b = rho0 has p_b = 0, so the pair term m (p_i/rho_i^2 + 0) grad W is the
    # boundary term with rho0 V_b = m (no extra 1/m); zero velocities and gravity leave only pressure
    P = _union_params()
    f = _cloud(300, 2)
    b = _wall(P.h)
    vol = np.full(len(b), P.mass / P.rest_dens, np.float32)
    rng = np.random.default_rng(3)
    dens = np.float32(P.rest_dens) * (1.0 + rng.uniform(-0.05, 0.1, len(f))).astype(np.float32)
    pres = (np.float32(P.gas_k) * (dens - np.float32(P.rest_dens))).astype(np.float32)
    zero = np.zeros_like(f)
    acc = np.zeros_like(f)
    orc.orc_sph_force_f(ptr(f), ptr(zero), ptr(dens), ptr(pres), len(f), C.byref(P), ptr(acc))
    orc.orc_boundary_force_f(ptr(f), len(f), ptr(b), ptr(vol), len(b), C.byref(P), ptr(dens), ptr(pres), ptr(acc))
    u = np.concatenate([f, b])
    du = np.concatenate([dens, np.full(len(b), np.float32(P.rest_dens))]).astype(np.float32)
    pu = np.zeros(len(u), np.float32)
    au, vu = np.zeros_like(u), np.zeros_like(u)
    orc.orc_sph_force_f(ptr(u), ptr(vu), ptr(du), ptr(pu), len(u), C.byref(P), ptr(au))
    scale = np.abs(au[:len(f), :3]).max()
    assert np.abs(acc[:, :3] - au[:len(f), :3]).max() <= 2e-5 * scale
    # the wall pushes the particles just above it upward (+z) when they are compressed
    near = (f[:, 2] < 0.33) & (pres > 0) & (np.abs(f[:, 0] - 0.5) < 0.2) & (np.abs(f[:, 1] - 0.5) < 0.2)
    acc_f = np.zeros_like(f)
    orc.orc_boundary_force_f(ptr(f), len(f), ptr(b), ptr(vol), len(b), C.byref(P), ptr(dens), ptr(pres), ptr(acc_f))
    assert near.sum() > 0 and np.all(acc_f[near, 2] > 0)


def test_multiplier_with_symmetric_boundary_equals_union(orc):
    # boundary particles in +-d pairs around fluid particle 0: their gradients cancel in the self term, so
    # lambda_0 with boundary = lambda_0 of the union, and every density equals the union's
    P = _union_params()
    f = _cloud(200, 4, 0.35, 0.65)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(12, 3))
    d *= (rng.uniform(0.2, 0.9, 12) * P.h / np.linalg.norm(d, axis=1))[:, None]
    c = f[0, :3].astype(np.float64)
    b = pos4(np.concatenate([c + d, c - d]))
    vol = np.full(len(b), P.mass / P.rest_dens, np.float32)
    dens, lam, lamb = np.zeros(len(f), np.float32), np.zeros(len(f), np.float32), np.zeros(len(b), np.float32)
    orc.orc_pbf_scaling_factor_wb_f(ptr(f), len(f), ptr(b), ptr(vol), len(b), C.byref(P), ptr(dens), ptr(lam), ptr(lamb))
    u = np.concatenate([f, b])
    du, lu = np.zeros(len(u), np.float32), np.zeros(len(u), np.float32)
    orc.orc_pbf_scaling_factor_f(ptr(u), len(u), C.byref(P), ptr(du), ptr(lu))
    np.testing.assert_allclose(dens, du[:len(f)], rtol=2e-6)
    assert lam[0] == pytest.approx(lu[0], rel=2e-5)
    # without the boundary the multiplier of particle 0 differs (the boundary terms are really there)
    d0, l0 = np.zeros(len(f), np.float32), np.zeros(len(f), np.float32)
    orc.orc_pbf_scaling_factor_f(ptr(f), len(f), C.byref(P), ptr(d0), ptr(l0))
    assert abs(l0[0] - lam[0]) > 1e-3 * abs(lam[0])


def test_boundary_multiplier_pair(orc):
    # one boundary particle b (volume V) and one fluid particle at distance r:
    # C_b = (m W(r) + rho0 V W(0)) / rho0 - 1, sd_b = |grad W(r)|^2 / rho0^2 (the self pair has r = 0)
    P = _union_params()
    h, m, r0 = P.h, P.mass, P.rest_dens
    r = 0.37 * h
    V = 2.5e-4
    f = pos4([[0.5 + r, 0.5, 0.5]])
    b = pos4([[0.5, 0.5, 0.5]])
    vol = np.array([V], np.float32)
    dens, lam, lamb = np.zeros(1, np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32)
    orc.orc_pbf_scaling_factor_wb_f(ptr(f), 1, ptr(b), ptr(vol), 1, C.byref(P), ptr(dens), ptr(lam), ptr(lamb))
    rr = float(np.float32(0.5 + r) - np.float32(0.5))
    Cb = (m * poly6(rr, h) + r0 * V * poly6(0.0, h)) / r0 - 1.0
    sdb = (spiky_grad_mag(rr, h) / r0) ** 2
    assert lamb[0] == pytest.approx(-Cb / (sdb + P.eps), rel=1e-5)
    # the fluid particle: rho = m W(0) + rho0 V W(r); its only fluid neighbour is itself (r = 0, no gradient), so
    # sd is the boundary gradient term alone, scaled by rho0 V / m
    rho = m * poly6(0.0, h) + r0 * V * poly6(rr, h)
    sd = (r0 * V / m * spiky_grad_mag(rr, h) / r0) ** 2
    assert dens[0] == pytest.approx(rho, rel=1e-6)
    assert lam[0] == pytest.approx(-(rho / r0 - 1.0) / (sd + P.eps), rel=1e-5)


def test_position_correction_equals_union(orc):
    # dp_i with boundary = dp_i of the union with lambda_b appended (V_b = m / rho0 makes s_corr identical)
    P = _union_params()
    f = _cloud(300, 6)
    b = _wall(P.h)
    rng = np.random.default_rng(7)
    lam = rng.uniform(-2e-5, 1e-5, len(f)).astype(np.float32)
    lamb = rng.uniform(-2e-5, 1e-5, len(b)).astype(np.float32)
    vol = np.full(len(b), P.mass / P.rest_dens, np.float32)
    dp = np.zeros_like(f)
    orc.orc_pbf_position_correction_wb_f(ptr(f), ptr(lam), len(f), ptr(b), ptr(vol), ptr(lamb), len(b), C.byref(P),
                                         np.float32(P.dt), ptr(dp))
    u = np.concatenate([f, b])
    lu = np.concatenate([lam, lamb]).astype(np.float32)
    dpu = np.zeros_like(u)
    orc.orc_pbf_position_correction_f(ptr(u), ptr(lu), len(u), C.byref(P), np.float32(P.dt), ptr(dpu))
    scale = np.abs(dpu[:len(f), :3]).max()
    assert scale > 0
    assert np.abs(dp[:, :3] - dpu[:len(f), :3]).max() <= 2e-5 * scale
