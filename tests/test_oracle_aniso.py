"""The CPU anisotropic-kernel path (oracle/aniso.cpp) against what fixes it independently: brute-force sums over
all particle pairs, numpy's symmetric eigen-decomposition and SVD, closed forms, and the isotropic lattice density."""
import ctypes as C

import numpy as np
import pytest

import oracle


def _sorted_grid(pos, origin, cell, gsize):
    """Cell hash (floor((p - o) / w) clamped, x fastest), stable sort, cell start / end (0xFFFFFFFF when empty)."""
    pos = np.asarray(pos, np.float32)
    g = np.floor((pos[:, :3] - np.float32(origin)) / np.float32(cell)).astype(np.int64)
    g = np.clip(g, 0, np.array(gsize) - 1)
    h = (g[:, 2] * gsize[1] + g[:, 1]) * gsize[0] + g[:, 0]
    order = np.argsort(h, kind="stable")
    hs = h[order]
    ncell = int(np.prod(gsize))
    cs = np.full(ncell, 0xFFFFFFFF, np.uint32)
    ce = np.full(ncell, 0xFFFFFFFF, np.uint32)
    u, first, cnt = np.unique(hs, return_index=True, return_counts=True)
    cs[u] = first
    ce[u] = first + cnt
    return np.ascontiguousarray(pos[order]), order.astype(np.uint32), cs, ce


def _cloud(n, seed, lo=0.0, hi=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 4), np.float32)
    p[:, :3] = rng.uniform(lo, hi, (n, 3))
    return p


def _run(pos, h, cell, lam=0.9, lattice=None, wpoly6=1.0, mass=1.0, h_eff=None):
    origin = np.array([-0.05, -0.05, -0.05], np.float32)
    gsize = np.ceil((1.1 / cell) * np.ones(3)).astype(np.int32)
    sp, sidx, cs, ce = _sorted_grid(pos, origin, cell, gsize)
    geo = np.concatenate([origin, np.full(3, cell, np.float32)])
    r = oracle.aniso_pipeline(sp, sidx, cs, ce, geo, gsize, lam, h, h_eff or h, mass, wpoly6, lattice)
    return r, (sp, sidx, cs, ce, geo, gsize)


@pytest.mark.parametrize("cell_over_h", [1.0, 0.6])
def test_weighted_mean_and_covariance_match_brute_force(cell_over_h):
    pos = _cloud(700, 1)
    h = 0.12
    r, _ = _run(pos, h, h * cell_over_h)
    P = pos[:, :3].astype(np.float64)
    d = np.linalg.norm(P[:, None, :] - P[None, :, :], axis=2)
    w = np.where(d < h, (1 - d / h) ** 3, 0.0)
    ws = w.sum(1)
    mean = (w @ P) / ws[:, None]
    assert np.allclose(r["wpos"][:, :3], mean, rtol=0, atol=2e-6)
    assert np.all(r["wpos"][:, 3] == 0)
    lam = np.float32(0.9)
    up = (1 - lam) * P + lam * mean
    assert np.allclose(r["up"][:, :3], up, atol=2e-6) and np.all(r["up"][:, 3] == 0)
    # covariance about the weighted mean, and the neighbour count (the particle itself included)
    dx = P[None, :, :] - mean[:, None, :]
    cov = np.einsum("ij,ijk,ijl->ikl", w, dx, dx) / ws[:, None, None]
    assert np.allclose(r["cmat"].reshape(-1, 3, 3), cov, rtol=1e-4, atol=1e-9)
    assert np.array_equal(r["posw"][:, 3], (d < h).sum(1).astype(np.float32))
    assert np.allclose(r["posw"][:, :3], mean, atol=2e-6)


def test_eigen_decomposition_matches_numpy():
    lib = oracle.load()
    rng = np.random.default_rng(3)
    for t in range(300):
        B = rng.normal(size=(3, 3))
        A = B @ B.T if t % 3 else B + B.T  # positive semi-definite, and indefinite
        if t % 7 == 0:
            A = np.diag(rng.uniform(0, 2, 3))  # already diagonal
        a = np.ascontiguousarray(A, np.float64).copy()
        v = np.zeros(9)
        lib.orc_sym_eigen3(oracle.ptr(a), oracle.ptr(v))
        lam, V = np.diag(a.reshape(3, 3)), v.reshape(3, 3)
        assert np.allclose(np.sort(lam), np.linalg.eigvalsh(A), rtol=1e-12, atol=1e-13 * np.abs(A).max())
        assert np.allclose(V @ V.T, np.eye(3), atol=1e-13)
        assert np.allclose(V @ np.diag(lam) @ V.T, A, atol=1e-12 * np.abs(A).max())


def _svd_stage(C9, counts):
    n = len(counts)
    posw = np.zeros((n, 4), np.float32)
    posw[:, 3] = counts
    eig = np.zeros((n, 3), np.float32)
    R = np.zeros((n, 9), np.float32)
    oracle.load().orc_aniso_svd(oracle.ptr(np.ascontiguousarray(C9, np.float32)), oracle.ptr(posw), n, oracle.ptr(eig), oracle.ptr(R))
    return eig, R.reshape(n, 3, 3)


def test_svd_stage_values_vectors_and_signs():
    rng = np.random.default_rng(5)
    n = 400
    Q = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    s = np.sort(rng.uniform(0.1, 1.0, (n, 3)), axis=1)[:, ::-1] * np.array([1.0, 0.8, 0.5])  # separated
    Cm = np.einsum("nij,nj,nkj->nik", Q, s, Q).astype(np.float32)
    counts = np.where(np.arange(n) % 5 == 0, 10.0, 25.0).astype(np.float32)
    eig, R = _svd_stage(Cm.reshape(n, 9), counts)
    sv = np.linalg.svd(Cm.astype(np.float64), compute_uv=False)  # descending
    exp = np.stack([sv[:, 0], np.maximum(sv[:, 1], sv[:, 0] / 4), np.maximum(sv[:, 2], sv[:, 0] / 4)], 1) * 1400
    big = counts > 10
    assert np.allclose(eig[big], exp[big], rtol=2e-5)
    assert np.all(eig[~big] == 0.5)
    # R's columns are the eigenvectors (numpy's SVD left vectors up to sign), orthonormal, reconstruct C
    U = np.linalg.svd(Cm.astype(np.float64))[0]
    dots = np.abs(np.einsum("nik,nik->nk", R.astype(np.float64), U))
    assert np.all(dots > 1 - 1e-4)
    rec = np.einsum("nij,nj,nkj->nik", R.astype(np.float64), sv, R.astype(np.float64))
    assert np.allclose(rec, Cm, atol=3e-6)
    # sign rule: no column has two or more negative entries (u_k = v_k for these matrices)
    assert np.all((R < 0).sum(axis=1) <= 1)


def test_svd_stage_closed_forms():
    # diagonal covariance diag(3, 1, 2): values descending with their axes, all entries non-negative
    eig, R = _svd_stage(np.diag([3.0, 1.0, 2.0]).reshape(1, 9), np.array([20.0], np.float32))
    assert np.array_equal(eig[0], np.float32([3 * 1400, 2 * 1400, 1 * 1400]))
    assert np.array_equal(R[0], np.float32([[1, 0, 0], [0, 0, 1], [0, 1, 0]]))
    # a flat neighbourhood: the two small values are lifted to a quarter of the largest
    eig, R = _svd_stage(np.diag([4.0, 0.1, 0.0]).reshape(1, 9), np.array([30.0], np.float32))
    assert np.array_equal(eig[0], np.float32([4 * 1400, 1400, 1400]))
    # isotropic C: equal values, identity axes
    eig, R = _svd_stage(np.eye(3).reshape(1, 9) * 0.5, np.array([11.0], np.float32))
    assert np.array_equal(eig[0], np.float32([700, 700, 700])) and np.array_equal(R[0], np.eye(3, dtype=np.float32))


def test_transform_is_normalised_inverse_shape():
    rng = np.random.default_rng(9)
    n = 200
    R = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0].astype(np.float32)
    s = rng.uniform(100.0, 2000.0, (n, 3)).astype(np.float32)
    G = np.zeros((n, 9), np.float32)
    oracle.load().orc_aniso_transform(oracle.ptr(s), oracle.ptr(np.ascontiguousarray(R.reshape(n, 9))), n, oracle.ptr(G))
    ref = np.einsum("nij,nj,nkj->nik", R.astype(np.float64), 1.0 / s, R.astype(np.float64))
    ref /= np.max(np.diagonal(ref, axis1=1, axis2=2), axis=1)[:, None, None]
    assert np.allclose(G.reshape(n, 3, 3), ref, atol=1e-5)
    assert np.allclose(np.max(np.diagonal(G.reshape(n, 3, 3), axis1=1, axis2=2), axis=1), 1.0, atol=1e-6)


def _lattice_density(stage, G9, nlat, lo, d, h, mass, wpoly6):
    sp, sidx, cs, ce, geo, gsize = stage
    out = np.zeros(int(np.prod(nlat)), np.float32)
    oracle.load().orc_aniso_grid_density(oracle.ptr(sp), oracle.ptr(sidx), oracle.ptr(cs), oracle.ptr(ce), len(sidx), oracle.ptr(geo),
                                         oracle.ptr(gsize), h, mass, wpoly6, oracle.ptr(np.ascontiguousarray(G9, np.float32)),
                                         *nlat, *lo, *d, oracle.ptr(out))
    return out


def test_lattice_density_matches_brute_force():
    pos = _cloud(600, 11, 0.2, 0.8)
    h, mass, wpoly6 = 0.1, 0.02, 315.0 / (64 * np.pi * 0.1**9)
    _, stage = _run(pos, h, 0.1)
    n = len(pos)
    rng = np.random.default_rng(2)
    # shapes with every eigenvalue in [0.4, 1]: the support |G r| <= h lies inside 2.5h, within the 3h cell box
    Q = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    ev = rng.uniform(0.4, 1.0, (n, 3))
    G = np.einsum("nij,nj,nkj->nik", Q, ev, Q).astype(np.float32)
    nlat, lo, d = (14, 12, 13), (0.15, 0.18, 0.16), (0.05, 0.06, 0.055)
    out = _lattice_density(stage, G.reshape(n, 9), nlat, lo, d, h, mass, wpoly6)
    z, y, x = np.meshgrid(*[lo[a] + np.arange(nlat[a]) * d[a] for a in (2, 1, 0)], indexing="ij")
    nodes = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    r = nodes[:, None, :] - pos[None, :, :3].astype(np.float64)
    rg = np.einsum("pij,npj->npi", G.astype(np.float64), r)
    rr = np.linalg.norm(rg, axis=2)
    det = np.linalg.det(G.astype(np.float64))
    ref = np.where(rr <= h, det[None, :] * mass * wpoly6 * (h * h - rr * rr) ** 3, 0.0).sum(1)
    assert (ref > 0).mean() > 0.5
    assert np.allclose(out, ref, rtol=2e-4, atol=1e-5 * ref.max())


def test_lattice_density_with_identity_shapes_is_the_isotropic_density():
    pos = _cloud(500, 4, 0.25, 0.75)
    h, mass, wpoly6 = 0.1, 0.02, 315.0 / (64 * np.pi * 0.1**9)
    _, stage = _run(pos, h, 0.1)
    n = len(pos)
    nlat, lo, d = (10, 11, 12), (0.2, 0.2, 0.2), (0.06, 0.055, 0.05)
    out = _lattice_density(stage, np.tile(np.eye(3, dtype=np.float32).ravel(), (n, 1)), nlat, lo, d, h, mass, wpoly6)
    z, y, x = np.meshgrid(*[lo[a] + np.arange(nlat[a]) * d[a] for a in (2, 1, 0)], indexing="ij")
    nodes = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    r2 = ((nodes[:, None, :] - pos[None, :, :3]) ** 2).sum(2)
    ref = np.where(r2 <= h * h, mass * wpoly6 * (h * h - r2) ** 3, 0.0).sum(1)
    assert np.allclose(out, ref, rtol=2e-4, atol=1e-5 * ref.max())


def test_sparse_neighbourhoods_take_the_isotropic_shape():
    # particles far apart: n <= 10 neighbours -> sigma = 0.5 on every axis -> G = I
    pos = _cloud(40, 6)
    r, _ = _run(pos, 0.05, 0.05)
    assert np.all(r["posw"][:, 3] <= 10)
    assert np.all(r["eigen"] == 0.5)
    assert np.allclose(r["gmat"].reshape(-1, 3, 3), np.eye(3), atol=1e-7)
