"""Pins the oracle's grid, neighbour search, WCSPH, PBF, XSPH and heat functions with checks that do not
retype its formulas: exact integer invariants, independent library routines (numpy's stable argsort,
scipy's KD-tree), conservation laws, linearisation of the PBF constraint, symplectic free fall and the
continuum limit of the SPH Laplacian."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle
from oracle import OrcParams, ptr


def make_params(h=0.1, mass=0.125, rest=998.29, gas_k=3.0, visc=0.0, g=(0.0, 0.0, 0.0), lo=-2.0, hi=2.0,
                eps=600.0, use_ap=0, iters=4, dt=1.0 / 120.0, radius=0.0, xsph=0.0) -> OrcParams:
    P = OrcParams()
    P.h, P.mass, P.rest_dens, P.gas_k, P.viscosity = h, mass, rest, gas_k, visc
    for k in range(3):
        P.gravity[k] = g[k]
        P.bmin[k], P.bmax[k] = lo, hi
        P.origin[k] = float(np.float32(lo - 0.01))
        P.cell[k] = float(np.float32(h))
        P.gsize[k] = int((hi - lo + 0.02) / h) + 1
    P.radius = radius
    P.eps, P.use_ap, P.ap_k, P.ap_n, P.ap_q = eps, use_ap, 0.1, 4.0, 0.2
    P.xsph_c, P.iterations, P.dt = xsph, iters, dt
    return P


def pos4(xyz) -> np.ndarray:
    xyz = np.asarray(xyz, np.float32)
    out = np.zeros((len(xyz), 4), np.float32)
    out[:, :3] = xyz
    return out


def poly6(r, h):  # Mueller et al. 2003, eq. (20)
    return np.where(r <= h, 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3, 0.0)


def lattice(nx, s, origin=(0.0, 0.0, 0.0), jitter=0.0, seed=0):
    g = np.stack(np.meshgrid(*[np.arange(nx)] * 3, indexing="ij"), -1).reshape(-1, 3) * s + np.asarray(origin)
    if jitter:
        g = g + np.random.default_rng(seed).uniform(-jitter, jitter, g.shape)
    return g


# ------------------------------------------------------------------------------------------ grid
def _hash(orc, p, P):
    n = len(p)
    hs = np.zeros(n, np.uint32)
    idx = np.zeros(n, np.uint32)
    orc.orc_calc_hash_f(ptr(p), None, n, C.byref(P), ptr(hs), ptr(idx))
    return hs, idx


def test_hash_cell_contains_point(orc):
    # origin 0 and a power-of-two cell: every cell face is exact in FP32, so the hash must name the cell
    # that contains the point (faces belong to the upper cell), clamped to the grid.
    P = make_params(h=0.125)
    for k in range(3):
        P.origin[k], P.cell[k], P.gsize[k] = 0.0, 0.125, 16
    rng = np.random.default_rng(1)
    x = rng.uniform(-0.3, 2.3, (5000, 3)).astype(np.float32)
    faces = (rng.integers(0, 17, (500, 3)) * 0.125).astype(np.float32)     # exactly on faces
    below = np.nextafter(faces, np.float32(-np.inf))                        # one ulp below a face
    p = pos4(np.concatenate([x, faces, below]))
    hs, idx = _hash(orc, p, P)
    cell = np.clip(np.floor(p[:, :3].astype(np.float64) / 0.125), 0, 15).astype(np.int64)
    expect = (cell[:, 2] * 16 + cell[:, 1]) * 16 + cell[:, 0]
    assert np.array_equal(hs.astype(np.int64), expect)
    assert np.array_equal(idx, np.arange(len(p), dtype=np.uint32))
    # deleted particles (attr -1) go to the 0x7FFFFFFF bucket
    attr = np.zeros(len(p), np.int32)
    attr[::7] = -1
    hs2 = np.zeros(len(p), np.uint32)
    orc.orc_calc_hash_f(ptr(p), ptr(attr), len(p), C.byref(P), ptr(hs2), ptr(idx))
    assert np.all(hs2[::7] == 0x7FFFFFFF)


def test_stable_sort_and_cell_bounds(orc):
    rng = np.random.default_rng(2)
    n, ncells = 20000, 997
    keys = rng.integers(0, ncells, n).astype(np.uint32)
    keys[rng.random(n) < 0.3] = 5  # long runs of ties
    hs = keys.copy()
    idx = np.arange(n, dtype=np.uint32)
    orc.orc_stable_sort(ptr(hs), ptr(idx), n)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(idx, order.astype(np.uint32))
    assert np.array_equal(hs, keys[order])
    start = np.zeros(ncells, np.uint32)
    end = np.zeros(ncells, np.uint32)
    orc.orc_cell_bounds(ptr(hs), n, ncells, ptr(start), ptr(end))
    counts = np.bincount(keys, minlength=ncells)
    empty = counts == 0
    assert np.all(start[empty] == 0xFFFFFFFF) and np.all(end[empty] == 0xFFFFFFFF)
    full = ~empty
    assert np.array_equal((end[full] - start[full]).astype(np.int64), counts[full])
    assert int(counts.sum()) == n
    for c in np.flatnonzero(full)[:50]:
        assert np.all(hs[start[c]:end[c]] == c)


def test_neighbours_match_kdtree(orc):
    rng = np.random.default_rng(3)
    h = 0.1
    x = rng.uniform(0, 0.6, (3000, 3)).astype(np.float32)
    d = np.linalg.norm(x[:, None, :].astype(np.float64) - x[None, :, :], axis=-1)
    assert np.min(np.abs(d - h)) > 1e-7  # no pair within FP32 rounding of the cut
    p = pos4(x)
    counts = np.zeros(len(p), np.int32)
    tot = orc.orc_neighbors_brute_f(ptr(p), len(p), h, 0, ptr(counts), None)
    tree = cKDTree(x.astype(np.float64))
    ref = np.array([len(v) for v in tree.query_ball_point(x.astype(np.float64), h)])
    assert np.array_equal(counts, ref) and tot == ref.sum()


def test_grid_density_matches_kdtree_sum(orc):
    # the density goes through the oracle's hashed grid query; the reference sum uses a KD-tree
    h, m = 0.1, 0.125
    x = lattice(10, 0.05, (0.2, 0.2, 0.2), jitter=0.01, seed=4).astype(np.float32)
    P = make_params(h=h, mass=m)
    p = pos4(x)
    dens = np.zeros(len(p), np.float32)
    pres = np.zeros(len(p), np.float32)
    orc.orc_sph_density_f(ptr(p), len(p), C.byref(P), ptr(dens), ptr(pres))
    x64 = x.astype(np.float64)
    tree = cKDTree(x64)
    ref = np.array([m * poly6(np.linalg.norm(x64[nb] - x64[i], axis=1), h).sum()
                    for i, nb in enumerate(tree.query_ball_point(x64, h))])
    np.testing.assert_allclose(dens, ref, rtol=2e-6)
    np.testing.assert_allclose(pres, 3.0 * (dens.astype(np.float64) - 998.29), rtol=1e-5, atol=1e-3)


# ------------------------------------------------------------------------------------------ WCSPH
def test_pressure_forces_conserve_momentum(orc):
    # with no gravity, no viscosity and equal masses the pair forces cancel: sum_i a_i = 0
    x = lattice(6, 0.04, (0.3, 0.3, 0.3), jitter=0.008, seed=5).astype(np.float32)
    P = make_params(h=0.1, mass=0.125, gas_k=3.0)
    p = pos4(x)
    n = len(p)
    dens = np.zeros(n, np.float32)
    pres = np.zeros(n, np.float32)
    acc = np.zeros((n, 4), np.float32)
    orc.orc_sph_density_f(ptr(p), n, C.byref(P), ptr(dens), ptr(pres))
    v = np.zeros((n, 4), np.float32)
    orc.orc_sph_force_f(ptr(p), ptr(v), ptr(dens), ptr(pres), n, C.byref(P), ptr(acc))
    s = acc[:, :3].astype(np.float64).sum(0)
    assert np.abs(s).max() < 1e-5 * np.abs(acc[:, :3]).astype(np.float64).sum()


def test_two_particle_pressure_and_viscosity(orc):
    # Pair along x at r = 0.4h.  Pressure: a_i = -m (p_i/rho_i^2 + p_j/rho_j^2) grad W_spiky(r_ij) with
    # grad W_spiky = -45/(pi h^6) (h - r)^2 r_hat (Monaghan's symmetric form with Mueller's kernel);
    # viscosity: a_i = mu m/rho_j (v_j - v_i) 45/(pi h^6)(h - r).
    h, m, r0, k, mu = 0.1, 0.7, 998.29, 3.0, 0.5
    r = 0.4 * h
    P = make_params(h=h, mass=m, gas_k=k, visc=mu, rest=r0)
    p = pos4([[0.5, 0.5, 0.5], [0.5 + r, 0.5, 0.5]])
    v = np.zeros((2, 4), np.float32)
    v[0, 1], v[1, 1] = 0.2, -0.3
    dens = np.zeros(2, np.float32)
    pres = np.zeros(2, np.float32)
    acc = np.zeros((2, 4), np.float32)
    orc.orc_sph_density_f(ptr(p), 2, C.byref(P), ptr(dens), ptr(pres))
    rr = float(np.float32(0.5 + r) - np.float32(0.5))
    rho = m * (poly6(0.0, h) + poly6(rr, h))
    np.testing.assert_allclose(dens, [rho, rho], rtol=1e-6)
    orc.orc_sph_force_f(ptr(p), ptr(v), ptr(dens), ptr(pres), 2, C.byref(P), ptr(acc))
    pr = k * (rho - r0)
    gmag = 45.0 / (math.pi * h ** 6) * (h - rr) ** 2
    ax0 = -m * (2 * pr / rho ** 2) * (-gmag) * (-1.0)  # r_hat of r_01 = x0 - x1 points along -x
    lap = 45.0 / (math.pi * h ** 6) * (h - rr)
    ay0 = mu * m / rho * (v[1, 1] - v[0, 1]) * lap
    np.testing.assert_allclose(acc[0, :3], [ax0, ay0, 0.0], rtol=2e-5, atol=1e-9)
    np.testing.assert_allclose(acc[1, :3], [-ax0, -ay0, 0.0], rtol=2e-5, atol=1e-9)
    assert ax0 < 0 if pr > 0 else ax0 > 0  # compressed pair repels: particle 0 sits at smaller x


def test_free_fall_symplectic_euler(orc):
    # one particle, gravity only: v_k = v0 + k g dt, x_k = x0 + k dt v0 + g dt^2 k (k + 1) / 2
    g, dt, K = -9.8, 1e-3, 50
    P = make_params(h=0.1, mass=0.1, g=(0.0, g, 0.0), gas_k=0.0, lo=-5.0, hi=5.0)
    p = pos4([[0.1, 2.0, -0.3]]).astype(np.float64)
    x0 = p[0, :3].copy()
    v = np.zeros((1, 4), np.float64)
    v[0, 0] = 0.25
    for _ in range(K):
        p, v, _ = oracle.sph_step(p, v, P, dt, double=True)
    np.testing.assert_allclose(v[0, :3], [0.25, K * g * dt, 0.0], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(p[0, :3], x0 + [K * dt * 0.25, g * dt * dt * K * (K + 1) / 2, 0.0],
                               rtol=1e-12, atol=1e-14)


def test_lattice_interior_feels_only_gravity(orc):
    # on a uniform lattice the pressure forces of an interior particle cancel by symmetry
    s = 0.05
    x = lattice(9, s, (0.3, 0.3, 0.3))
    P = make_params(h=0.1, mass=0.125, gas_k=3.0, g=(0.0, -9.8, 0.0))
    p = pos4(x)
    n = len(p)
    dens = np.zeros(n, np.float32)
    pres = np.zeros(n, np.float32)
    acc = np.zeros((n, 4), np.float32)
    orc.orc_sph_density_f(ptr(p), n, C.byref(P), ptr(dens), ptr(pres))
    vel = np.zeros((n, 4), np.float32)
    orc.orc_sph_force_f(ptr(p), ptr(vel), ptr(dens), ptr(pres), n, C.byref(P), ptr(acc))
    centre = np.argmin(np.linalg.norm(x - x.mean(0), axis=1))
    np.testing.assert_allclose(acc[centre, :3], [0.0, -9.8, 0.0], atol=1e-3)


# ------------------------------------------------------------------------------------------ PBF
def _lambda(orc, p, P):
    n = len(p)
    d = np.zeros(n, np.float32)
    lam = np.zeros(n, np.float32)
    orc.orc_pbf_scaling_factor_f(ptr(p), n, C.byref(P), ptr(d), ptr(lam))
    return d, lam


def _dp(orc, p, lam, P, dt):
    dp = np.zeros((len(p), 4), np.float32)
    orc.orc_pbf_position_correction_f(ptr(p), ptr(lam), len(p), C.byref(P), dt, ptr(dp))
    return dp


def test_pbf_lambda_sign_and_zero_at_rest(orc):
    x = lattice(6, 0.045, (0.3, 0.3, 0.3), jitter=0.005, seed=6).astype(np.float32)
    P = make_params(h=0.1, mass=0.125)
    d, lam = _lambda(orc, pos4(x), P)
    C_ = d.astype(np.float64) / 998.29 - 1.0
    nz = np.abs(C_) > 1e-4
    assert np.all(np.sign(lam[nz]) == -np.sign(C_[nz]))
    # a lone particle whose self density equals rho0 has C = 0, lambda = 0
    m0 = 998.29 / poly6(0.0, 0.1)
    P0 = make_params(h=0.1, mass=float(np.float32(m0)))
    d0, l0 = _lambda(orc, pos4([[0.5, 0.5, 0.5]]), P0)
    assert abs(d0[0] / 998.29 - 1.0) < 1e-6 and abs(l0[0]) < 1e-9


@pytest.mark.parametrize("grad_mass", [0, 1])
@pytest.mark.parametrize("rest_frac", [0.45, 0.55, 0.7])
def test_pbf_pair_correction_is_linearised_decay(orc, rest_frac, grad_mass):
    # Pair, eps = 0, no s_corr.  C = (m/rho0) sum W_poly6 while the correction uses the spiky gradient
    # g = c grad W_spiky / rho0 (c = 1 in the reference's form, c = m when the gradient carries the mass):
    # dp_i = (lambda_i + lambda_j) g, lambda = -C / (2 |g|^2) (self term + pair term).
    # To first order the constraint then changes by
    #   dC = (m/rho0) grad W_poly6 . (dp_i - dp_j) = -2 (m / c) C |grad W_poly6| / |grad W_spiky|,
    # i.e. the reference's factor carries m, the mass-consistent one does not.
    # A dropped self term in the denominator, a sign error or a missing 1/rho0 changes the factor.
    h, r0 = 0.1, 998.29
    rest_r = rest_frac * h
    m = float(np.float32(r0 / (poly6(0.0, h) + poly6(rest_r, h))))  # rest_r is the rest distance
    P = make_params(h=h, mass=m, eps=0.0)
    P.grad_mass = grad_mass
    for delta in (3e-3, 1e-3):
        r = rest_r * (1 - delta)
        p = pos4([[0.5, 0.5, 0.5], [0.5 + r, 0.5, 0.5]])
        d, lam = _lambda(orc, p, P)
        C0 = d[0] / r0 - 1.0
        dp = _dp(orc, p, lam, P, 1.0 / 120.0)
        np.testing.assert_allclose(dp[0, :3] + dp[1, :3], 0.0, atol=1e-12)
        assert dp[0, 0] < 0 < dp[1, 0]  # a compressed pair is pushed apart
        p2 = p.copy()
        p2[:, :3] += dp[:, :3]
        d2, _ = _lambda(orc, p2, P)
        C1 = d2[0] / r0 - 1.0
        rr = float(np.float32(0.5 + r) - np.float32(0.5))
        g_poly6 = 945.0 / (32.0 * math.pi * h ** 9) * rr * (h * h - rr * rr) ** 2
        g_spiky = 45.0 / (math.pi * h ** 6) * (h - rr) ** 2
        ratio = 1.0 - 2.0 * (1.0 if grad_mass else m) * g_poly6 / g_spiky
        # second-order remainder: the mass-consistent correction is about 1/m times larger, so its remainder is too
        bound = (4.0 * delta + 1e-3) * (3.0 if grad_mass else 1.0)
        assert abs(C1 / C0 - ratio) < bound, (C1 / C0, ratio)


def test_pbf_correction_is_momentum_neutral(orc):
    # pair terms are antisymmetric (s_corr is symmetric in i, j): sum_i dp_i = 0
    x = lattice(6, 0.042, (0.3, 0.3, 0.3), jitter=0.006, seed=7).astype(np.float32)
    for ap in (0, 1):
        P = make_params(h=0.1, mass=0.125, use_ap=ap)
        p = pos4(x)
        _, lam = _lambda(orc, p, P)
        dp = _dp(orc, p, lam, P, 1.0 / 120.0)
        s = dp[:, :3].astype(np.float64).sum(0)
        assert np.abs(s).max() < 1e-5 * np.abs(dp[:, :3]).astype(np.float64).sum()


def test_pbf_iterations_reduce_compression(orc):
    # a compressed block without walls or gravity: the mean density error falls at every iteration
    x = lattice(8, 0.042, (0.3, 0.3, 0.3), jitter=0.004, seed=8).astype(np.float32)
    P = make_params(h=0.1, mass=0.125, iters=6, eps=600.0)
    _, _, _, st = oracle.pbf_step(pos4(x), np.zeros((len(x), 4), np.float32), P)
    e = [st.mean_abs[k] for k in range(6)]
    assert all(b < a for a, b in zip(e, e[1:])), e
    assert st.max_err[5] < st.max_err[0]


def test_xsph_pair_and_uniform_field(orc):
    h, m, c = 0.1, 0.125, 0.3
    P = make_params(h=h, mass=m)
    x = lattice(5, 0.05, (0.3, 0.3, 0.3), jitter=0.005, seed=9).astype(np.float32)
    p = pos4(x)
    n = len(p)
    dens = np.zeros(n, np.float32)
    orc.orc_pbf_density_f(ptr(p), n, C.byref(P), ptr(dens))
    v = np.zeros((n, 4), np.float32)
    v[:, :3] = [0.3, -1.0, 2.0]
    out = np.zeros((n, 4), np.float32)
    orc.orc_pbf_xsph_f(ptr(p), ptr(v), ptr(dens), n, C.byref(P), c, ptr(out))
    np.testing.assert_allclose(out[:, :3], v[:, :3], rtol=1e-6)  # a uniform field is left alone
    # pair: v_i' = v_i + c m/rho_j (v_j - v_i) W(r)
    r = 0.6 * h
    p2 = pos4([[0.5, 0.5, 0.5], [0.5, 0.5 + r, 0.5]])
    v2 = np.zeros((2, 4), np.float32)
    v2[0, 0], v2[1, 0] = 1.0, -1.0
    d2 = np.zeros(2, np.float32)
    orc.orc_pbf_density_f(ptr(p2), 2, C.byref(P), ptr(d2))
    o2 = np.zeros((2, 4), np.float32)
    orc.orc_pbf_xsph_f(ptr(p2), ptr(v2), ptr(d2), 2, C.byref(P), c, ptr(o2))
    rr = float(np.float32(0.5 + r) - np.float32(0.5))
    expect = 1.0 + c * m / float(d2[1]) * (-2.0) * poly6(rr, h)
    assert o2[0, 0] == pytest.approx(expect, rel=1e-6)
    assert o2[0, 0] + o2[1, 0] == pytest.approx(0.0, abs=1e-6)  # equal densities: momentum kept


# ------------------------------------------------------------------------------------------ heat
def _heat(orc, p, T, P, kappa, dt):
    n = len(p)
    dens = np.zeros(n, np.float32)
    orc.orc_pbf_density_f(ptr(p), n, C.byref(P), ptr(dens))
    out = np.zeros(n, np.float32)
    T32 = np.ascontiguousarray(T, np.float32)
    orc.orc_heat_diffuse_f(ptr(p), ptr(dens), ptr(T32), n, C.byref(P), kappa, dt, ptr(out))
    return dens, out


def test_heat_conserves_volume_weighted_energy(orc):
    x = lattice(7, 0.045, (0.3, 0.3, 0.3), jitter=0.008, seed=10).astype(np.float32)
    P = make_params(h=0.1, mass=0.125)
    T = np.random.default_rng(11).uniform(0.0, 20.0, len(x)).astype(np.float32)
    dens, T2 = _heat(orc, pos4(x), T, P, 1e-4, 1e-2)
    w = 0.125 / dens.astype(np.float64)
    assert np.dot(w, T2) == pytest.approx(np.dot(w, T), rel=1e-6)
    assert T2.min() >= T.min() - 1e-3 and T2.max() <= T.max() + 1e-3
    _, Tc = _heat(orc, pos4(x), np.full(len(x), 7.5, np.float32), P, 1e-4, 1e-2)
    np.testing.assert_allclose(Tc, 7.5, rtol=0, atol=1e-6)


def test_heat_laplacian_of_quadratic_field(orc):
    # T = |x - c|^2 has Laplacian 6.  sum_j V_j (T_j - T_i) 45/(pi h^6)(h - r) tends to it on a fine lattice:
    # int r^2 45/(pi h^6)(h - r) dV = 6 over the ball of radius h.
    h, s = 0.1, 0.0125
    x = lattice(25, s, (0.3, 0.3, 0.3))
    c = x.mean(0)
    P = make_params(h=h)
    m = float(np.float32(998.29 * s ** 3))
    P.mass = m
    T = (np.sum((x - c) ** 2, axis=1)).astype(np.float32)
    kappa, dt = 1.0, 1e-3
    dens, T2 = _heat(orc, pos4(x), T, P, kappa, dt)
    centre = np.argmin(np.linalg.norm(x - c, axis=1))
    lap = (float(T2[centre]) - float(T[centre])) / (kappa * dt)
    assert lap == pytest.approx(6.0, rel=0.05)


# ------------------------------------------------------------------------------------------ walls
def test_wall_projection_is_the_clamp_to_the_shrunk_box(orc):
    # The wall contact (rxSolidBox distance with the particle radius) must move every particle to the nearest
    # point of the box shrunk by the radius -- for points beyond one face, beyond two (edges) and beyond three
    # (corners), and for points inside the box but closer than r to a face -- and leave the velocity alone.
    # The nearest point of a box is the per-axis clamp.
    lo, hi, r, dt = -1.0, 1.0, 0.07, 0.01
    P = make_params(h=0.1, lo=lo, hi=hi, radius=r)
    rng = np.random.default_rng(11)
    x0 = rng.uniform(lo - 0.4, hi + 0.4, (6000, 3))
    v = rng.normal(0.0, 2.0, (6000, 3))
    a = rng.normal(0.0, 5.0, (6000, 3))
    p, vv, aa = pos4(x0), pos4(v), pos4(a)
    n = len(p)
    out_p, out_v = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
    orc.orc_integrate_polygon_f(ptr(p), ptr(vv), ptr(aa), None, n, C.byref(P), None, None, 0, None, None, None, 0.0,
                                np.float32(dt), ptr(out_p), ptr(out_v))
    v1 = vv[:, :3].astype(np.float64) + np.float64(np.float32(dt)) * aa[:, :3]
    x1 = p[:, :3].astype(np.float64) + np.float64(np.float32(dt)) * v1
    expect = np.clip(x1, lo + r, hi - r)
    outside = ((x1 < lo + r) | (x1 > hi - r)).sum(1)
    assert all((outside == k).sum() > 50 for k in (0, 1, 2, 3))  # inside, faces, edges and corners all occur
    np.testing.assert_allclose(out_p[:, :3], expect, rtol=0, atol=2e-6)
    np.testing.assert_allclose(out_v[:, :3], v1, rtol=1e-6, atol=1e-6)  # the wall does not touch the velocity


# ------------------------------------------------------------------------------------------ PBF external force
def test_pbf_external_force_pair_closed_form(orc):
    # Pair along x at r = 0.35h with given densities rho_i != rho_j: the viscous acceleration of the PBF
    # prediction divides by the particle's own density (rx_sph_pbd.cpp:676),
    #   a_i = mu m / rho_i  45/(pi h^6) (h - r) (v_j - v_i) + g,
    # so swapping in rho_j (the WCSPH form), a sign error or a missing mu/m changes the result.  The viscous
    # pair force also removes kinetic energy: (v_i - v_j) . (rho_i a_i - rho_j a_j) <= 0 with gravity taken out.
    h, m, mu = 0.1, 0.3, 0.8
    g = (0.0, -9.8, 0.0)
    P = make_params(h=h, mass=m, visc=mu, g=g)
    r = 0.35 * h
    p = pos4([[0.2, 0.1, -0.3], [0.2 + r, 0.1, -0.3]])
    v = pos4([[0.4, -0.2, 0.1], [-0.1, 0.3, 0.25]])
    dens = np.array([900.0, 1100.0], np.float32)
    acc = np.zeros((2, 4), np.float32)
    orc.orc_pbf_ext_force_f(ptr(p), ptr(v), ptr(dens), 2, C.byref(P), ptr(acc))
    rr = float(np.float32(0.2 + r) - np.float32(0.2))
    lap = 45.0 / (math.pi * h ** 6) * (h - rr)
    dv = v[1, :3].astype(np.float64) - v[0, :3]
    e0 = mu * m / 900.0 * lap * dv + np.array(g)
    e1 = mu * m / 1100.0 * lap * (-dv) + np.array(g)
    np.testing.assert_allclose(acc[0, :3], e0, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(acc[1, :3], e1, rtol=1e-6, atol=1e-6)
    f0 = dens[0] * (acc[0, :3].astype(np.float64) - g)
    f1 = dens[1] * (acc[1, :3].astype(np.float64) - g)
    np.testing.assert_allclose(f0 + f1, 0.0, atol=1e-6 * np.abs(f0).max())  # equal and opposite viscous forces
    assert float(-dv @ (f0 - f1)) <= 0.0


# ------------------------------------------------------------------------------------------ FP64 PBF step
def test_pbf_step_fp64_matches_fp32_to_rounding(orc):
    # The FP64 drift oracle runs the same step as the FP32 one; after one step (4 iterations, s_corr, XSPH,
    # walls and gravity) the two agree to FP32 rounding of the positions and of the velocity quotient.
    from fltk_sph_ice_amd import scene

    sc = scene.pbf_dam_break(10, 10, 10)
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    p32, v32, d32, st32 = oracle.pbf_step(pos, vel, P)
    p64, v64, d64, st64 = oracle.pbf_step(pos.astype(np.float64), vel.astype(np.float64), P, double=True)
    scale = np.abs(p64[:, :3]).max()
    perr = np.abs(p32[:, :3] - p64[:, :3]).max() / scale
    assert perr < 2e-6, perr
    # v = (x* - x)/dt amplifies position rounding by 1/dt
    verr = np.abs(v32[:, :3] - v64[:, :3]).max()
    assert verr < 4.0 * 1.2e-7 * scale / sc.dt, verr
    np.testing.assert_allclose(d32, d64, rtol=2e-6)
    for k in range(sc.iterations):
        assert st32.max_err[k] == pytest.approx(st64.max_err[k], rel=1e-3, abs=1e-6)
    assert st64.iterations == sc.iterations


# ------------------------------------------------------------------------------------------ first-correction factor and cap
def test_first_correction_factor_scales_and_cap_bounds_the_correction(orc):
    # One correction, no s_corr, no XSPH: the correction is sum_j (lambda_i + lambda_j) grad W, linear in the
    # multipliers, so a factor s on them gives s times the correction (to FP64 rounding of the FP32 sums);
    # a cap holds every particle's correction to that length and leaves the shorter ones as they were.
    from fltk_sph_ice_amd import scene

    sc = scene.pbf_dam_break(10, 10, 10)
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    P.use_ap = 0
    P.xsph_c = 0.0
    P.iterations = 0
    pred, _, _, _ = oracle.pbf_step(pos, vel, P, double=True)  # x*: the prediction alone
    P.iterations = 1
    one, _, _, _ = oracle.pbf_step(pos, vel, P, double=True)
    P.first_scale = 2.0
    two, _, _, _ = oracle.pbf_step(pos, vel, P, double=True)
    d1 = one[:, :3] - pred[:, :3]
    d2 = two[:, :3] - pred[:, :3]
    inside = np.all((two[:, :3] > sc.box_min[0] + 0.06) & (two[:, :3] < np.asarray(sc.box_max) - 0.06), axis=1)  # no wall clamp
    assert inside.sum() > 100
    np.testing.assert_allclose(d2[inside], 2.0 * d1[inside], rtol=1e-6, atol=1e-12)
    P.first_scale = 1.0
    cap = 0.5 * float(np.linalg.norm(d1, axis=1).max())
    P.first_dmax = cap
    capped, _, _, _ = oracle.pbf_step(pos, vel, P, double=True)
    dc = capped[:, :3] - pred[:, :3]
    lc, l1 = np.linalg.norm(dc, axis=1), np.linalg.norm(d1, axis=1)
    assert lc.max() <= cap * (1 + 1e-6)
    short = l1 < cap * 0.999
    assert np.array_equal(capped[short], one[short]) and (~short).sum() > 0
    longer = ~short & inside
    np.testing.assert_allclose(dc[longer], d1[longer] * (cap / l1[longer])[:, None], rtol=1e-6)
