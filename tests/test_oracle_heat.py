"""Pins the oracle's restatement of LaunchHeatTransferGPU (HeatTransfer.cu:44-204) and of the 30-slot
neighbour / surface detection (rx_sph_kernel.cu:1134-1290) with hand-worked cases and mathematical facts."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.spatial import cKDTree

from oracle import OrcParams, ptr


def heat(orc, heats, temps, surf, neighs, dens, phase, melt=None, radius=2.0, air=10.0, ht=0.5, td=0.3, dt=0.01):
    n = len(temps)
    heats = np.array(heats, np.float32)
    temps = np.array(temps, np.float32)
    dtemps = np.zeros(n, np.float32)
    phase = np.array(phase, np.int32)
    changed = np.zeros(n, np.int32)
    surf = np.array(surf, np.int32)
    neighs = np.array(neighs, np.int32).reshape(n, 30)
    dens = np.array(dens, np.float32)
    m = None if melt is None else np.ascontiguousarray(melt, np.int32)
    orc.orc_heat_transfer_gpu_f(ptr(heats), ptr(temps), ptr(dtemps), ptr(surf), ptr(neighs), ptr(dens), ptr(phase),
                                ptr(changed), None if m is None else ptr(m), radius, air, ht, td, dt, n)
    return heats, temps, dtemps, phase, changed


def lists(n, pairs):
    nb = -np.ones((n, 30), np.int32)
    nb[:, 29] = 0
    for i, js in pairs.items():
        nb[i, :len(js)] = js
        nb[i, 29] = len(js)
    return nb


def test_spline_normalisation_2d():
    # the cubic B-spline of Monaghan (1992) with sigma = 10 / (7 pi h^2) integrates to 1 over the plane
    h = 0.7
    s = 10.0 / (7.0 * math.pi * h * h)

    def w(r):
        q = r / h
        return s * (1 - 1.5 * q * q + 0.75 * q ** 3) if q < 1 else (s * 0.25 * (2 - q) ** 3 if q < 2 else 0.0)

    total = quad(lambda r: w(r) * 2 * math.pi * r, 0, 2 * h, points=[h])[0]
    assert total == pytest.approx(1.0, rel=1e-10)


def test_sensible_heat_and_diffusion_pair(orc):
    # Two ice particles below the melt point, listed as each other's neighbours (plus themselves).
    # dT_i = td * dt (T_j - T_i) / rho_j * W(1; r) with W(1; r=2) = sigma (1 - 1.5/4 + 0.75/8), sigma = 10/(7 pi 4).
    r = 2.0
    w = 10.0 / (7.0 * math.pi * r * r) * (1 - 1.5 * 0.25 + 0.75 * 0.125)
    nb = lists(2, {0: [0, 1], 1: [1, 0]})
    h, t, dT, ph, ch = heat(orc, [0, 2.1], [100.0, 40.0], [0, 0], nb, [2.0, 4.0], [-2, -2], air=0.0, radius=r)
    d0 = 0.3 * 0.01 * (40.0 - 100.0) / 4.0 * w
    d1 = 0.3 * 0.01 * (100.0 - 40.0) / 2.0 * w
    np.testing.assert_allclose(dT, [d0, d1], rtol=1e-6)
    # sensible update: T += heats / c (c = 2.1 for T <= 250) then T += dT; heats reset
    np.testing.assert_allclose(t, [100.0 + d0, 40.0 + 1.0 + d1], rtol=1e-6)
    assert np.all(h == 0) and np.all(ph == -2) and np.all(ch == 0)


def test_air_exchange_of_surface_particles(orc):
    # surface particle with 5 neighbours: area (20 - 5) / 20 = 0.75; q = ht (T_air - T) area, then T += q / 2.1
    nb = lists(2, {})
    h, t, _, _, _ = heat(orc, [0, 0], [100.0, 100.0], [5, 0], nb, [1, 1], [-2, -2], air=20.0, ht=0.5)
    q = 0.5 * (20.0 - 100.0) * 0.75
    assert t[0] == pytest.approx(100.0 + q / 2.1, rel=1e-6)
    assert t[1] == 100.0
    # 25 neighbours: no exposed area
    h, t, _, _, _ = heat(orc, [0], [100.0], [25], lists(1, {}), [1], [-2], air=20.0)
    assert t[0] == 100.0


def test_phase_machine(orc):
    nb = lists(6, {})
    temps = [260.0, 240.0, 250.0, 250.0, 600.0, 250.0]
    heats = [0.0, 0.0, 170.0, -1.0, 0.0, 100.0]
    phase = [-2, 2, -1, 1, 2, -1]
    h, t, _, ph, ch = heat(orc, heats, temps, [0] * 6, nb, [1] * 6, phase, air=0.0)
    # 0: ice above 250 -> melting, T = 250, heat 0
    assert (ph[0], t[0], h[0], ch[0]) == (-1, 250.0, 0.0, 0)
    # 1: water below 250 -> freezing with the latent heat stored
    assert (ph[1], t[1], h[1], ch[1]) == (1, 250.0, 165.0, 0)
    # 2: melting with more than the latent heat -> water at 251, flagged
    assert (ph[2], t[2], h[2], ch[2]) == (2, 251.0, 0.0, 1)
    # 3: freezing with negative heat -> ice at 249, flagged
    assert (ph[3], t[3], h[3], ch[3]) == (-2, 249.0, 0.0, 1)
    # 4: water hotter than TEMP_MAX is clamped
    assert (ph[4], t[4]) == (2, 500.0)
    # 5: melting, heat within the latent range: stays melting
    assert (ph[5], t[5], h[5]) == (-1, 250.0, 100.0)


def test_melt_entries(orc):
    n = 8
    melt = -np.ones(1000, np.int32)
    melt[0] = 3          # slot 0 holds the count and is never applied
    melt[1], melt[2], melt[3] = 5, 5, 2
    h, t, _, _, _ = heat(orc, [0] * n, [0.0] * n, [0] * n, lists(n, {}), [1] * n, [-2] * n, melt=melt, air=0.0)
    # 2 x 300 to particle 5, 300 to particle 2 (converted to temperature with c = 2.1 and clamped to 500)
    assert t[5] == pytest.approx(min(600.0 / 2.1, 500.0)) and t[2] == pytest.approx(300.0 / 2.1)
    assert t[3] == 0.0 and t[0] == 0.0


def _params(h, lo, hi):
    P = OrcParams()
    P.h = h
    for k in range(3):
        P.origin[k] = lo
        P.cell[k] = h
        P.gsize[k] = int((hi - lo) / h) + 1
    return P


def test_detect_neighbours_sets_and_order(orc):
    rng = np.random.default_rng(4)
    h = 0.1
    x = rng.uniform(0.0, 0.5, (1500, 3)).astype(np.float32)
    p = np.zeros((len(x), 4), np.float32)
    p[:, :3] = x
    P = _params(h, 0.0, 0.5)
    out = np.zeros((len(x), 30), np.int32)
    orc.orc_detect_neighbors_f(ptr(p), len(x), C.byref(P), ptr(out))
    tree = cKDTree(x.astype(np.float64))
    full = tree.query_ball_point(x.astype(np.float64), h)
    g = np.floor(x / np.float32(h)).astype(np.int64)
    hsh = (g[:, 2] * P.gsize[1] + g[:, 1]) * P.gsize[0] + g[:, 0]
    for i in range(len(x)):
        cnt = out[i, 29]
        assert cnt == min(len(full[i]), 30)
        listed = out[i, :min(cnt, 29)]
        expect = sorted(full[i], key=lambda j: (hsh[j], j))[:min(cnt, 29)]
        assert list(listed) == expect
        assert np.all(out[i, cnt:29] == -1)


def test_detect_surface_rules(orc):
    # an isolated pair: count 2 <= 3 -> surface flag = count; a lattice interior: centre of mass at the
    # particle -> not surface
    s = 0.05
    g = np.stack(np.meshgrid(*[np.arange(5)] * 3, indexing="ij"), -1).reshape(-1, 3) * s
    p = np.zeros((len(g) + 2, 4), np.float32)
    p[:len(g), :3] = g
    p[len(g):, :3] = [[3.0, 3.0, 3.0], [3.05, 3.0, 3.0]]
    P = _params(0.1, 0.0, 3.2)
    nb = np.zeros((len(p), 30), np.int32)
    orc.orc_detect_neighbors_f(ptr(p), len(p), C.byref(P), ptr(nb))
    surf = np.zeros(len(p), np.int32)
    orc.orc_detect_surface_f(ptr(nb), ptr(surf), ptr(p), len(p), 1.0, 0.05)
    centre = 2 * 25 + 2 * 5 + 2
    assert surf[centre] == 0
    assert surf[-1] == 2 and surf[-2] == 2
    corner = 0
    assert surf[corner] == nb[corner, 29]


def _therm(T, heat=None, phase=None):
    n = len(T)
    th = np.zeros((n, 4), np.float32)
    th[:, 0] = T
    th[:, 1] = 0 if heat is None else heat
    th[:, 2] = np.where(np.asarray(T) < 250, -2, 2) if phase is None else phase
    return th


def test_cpu_heat_step_pair_and_conservation(orc):
    h = 0.1
    P = _params(h, 0.0, 1.0)
    # pair at r = 0.5h: dT_0 = td dt (T1 - T0) / rho1 W(0.5h), W = 10/(7 pi h^2)(1 - 1.5/4 + 0.75/8)
    p = np.zeros((2, 4), np.float32)
    p[:, :3] = [[0.5, 0.5, 0.5], [0.55, 0.5, 0.5]]
    dens = np.array([900.0, 1000.0], np.float32)
    th = _therm([240.0, 260.0], phase=[-2, 2])
    orc.orc_heat_cpu_f(ptr(p), ptr(dens), ptr(th), 2, C.byref(P), 50.0, 0.01)
    r = float(np.float32(0.55) - np.float32(0.5)) / h
    w = 10.0 / (7.0 * math.pi * h * h) * (1 - 1.5 * r * r + 0.75 * r ** 3)
    d0 = 50.0 * 0.01 * 20.0 / 1000.0 * w
    d1 = 50.0 * 0.01 * -20.0 / 900.0 * w
    assert th[0, 0] == pytest.approx(240.0 + d0, rel=1e-6) and th[1, 0] == pytest.approx(260.0 + d1, rel=1e-6)
    # many particles, uniform density, all sensible: sum of rho-weighted changes vanishes (antisymmetric pairs)
    rng = np.random.default_rng(3)
    x = rng.uniform(0.2, 0.5, (600, 3)).astype(np.float32)
    p = np.zeros((600, 4), np.float32)
    p[:, :3] = x
    dens = np.full(600, 1000.0, np.float32)
    T0 = rng.uniform(100.0, 240.0, 600).astype(np.float32)
    th = _therm(T0, phase=np.full(600, -2))
    orc.orc_heat_cpu_f(ptr(p), ptr(dens), ptr(th), 600, C.byref(P), 1.0, 0.01)
    assert float(np.sum(th[:, 0].astype(np.float64) - T0)) == pytest.approx(0.0, abs=1e-3 * np.abs(th[:, 0] - T0).sum())
    # a constant field stays constant
    th = _therm(np.full(600, 120.0, np.float32), phase=np.full(600, -2))
    orc.orc_heat_cpu_f(ptr(p), ptr(dens), ptr(th), 600, C.byref(P), 1.0, 0.01)
    assert np.all(th[:, 0] == 120.0)
