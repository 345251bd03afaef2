"""Pins the oracle's shape matching (ShapeMatching.h / ShapeMatching.cpp restated) with closed forms."""
import ctypes as C

import numpy as np
import pytest

import oracle


def _rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_jacobi_matches_eigh(orc):
    rng = np.random.default_rng(0)
    for _ in range(50):
        m = rng.normal(size=(3, 3))
        a = (m @ m.T + 0.1 * np.eye(3)).astype(np.float64)
        w_ref = np.sort(np.linalg.eigvalsh(a))
        aa = np.ascontiguousarray(a.copy())
        v = np.ascontiguousarray(np.eye(3))
        orc.orc_jacobi_eigen3(oracle.ptr(aa), oracle.ptr(v), 1e-12, 100)
        np.testing.assert_allclose(np.sort(np.diag(aa)), w_ref, rtol=1e-9, atol=1e-12)
        # columns of v are the eigenvectors: a v = v diag
        np.testing.assert_allclose(a @ v, v @ np.diag(np.diag(aa)), atol=1e-8)


def test_polar_decomposition_properties(orc):
    rng = np.random.default_rng(1)
    for _ in range(50):
        A = rng.normal(size=(3, 3))
        if np.linalg.det(A) < 0:
            A[:, 0] = -A[:, 0]
        A = np.ascontiguousarray(A)
        R = np.zeros((3, 3))
        S = np.zeros((3, 3))
        orc.orc_polar_decomposition(oracle.ptr(A), oracle.ptr(R), oracle.ptr(S))
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-7)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-7)
        np.testing.assert_allclose(S, S.T, atol=1e-7)
        assert np.all(np.linalg.eigvalsh(0.5 * (S + S.T)) > -1e-9)
        np.testing.assert_allclose(R @ S, A, atol=1e-7)
        # independent route: R = U V^T from the SVD
        U, _, Vt = np.linalg.svd(A)
        np.testing.assert_allclose(R, U @ Vt, atol=1e-6)


def _lattice(n=5, s=0.1):
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3) * s
    p = np.zeros((len(g), 4))
    p[:, :3] = g
    return p


def _clusters(rest, radius):
    d = np.linalg.norm(rest[:, None, :3] - rest[None, :, :3], axis=-1)
    start = [0]
    mem = []
    for i in range(len(rest)):
        m = np.nonzero(d[i] <= radius)[0]
        mem.extend(m.tolist())
        start.append(len(mem))
    return np.array(start, np.int32), np.array(mem, np.int32)


def _params(alpha=1.0, beta=0.0, iters=1, g=(0, 0, 0), dt=0.01):
    S = oracle.OrcSmParams()
    S.alpha, S.beta, S.dt = alpha, beta, dt
    for k in range(3):
        S.gravity[k] = g[k]
        S.bmin[k] = -100.0
        S.bmax[k] = 100.0
    S.iterations = iters
    S.volume_conservation = 1
    S.restitution = 0.9
    return S


@pytest.mark.parametrize("beta", [0.0, 0.3, 1.0])
def test_rigid_motion_is_a_fixed_point(orc, beta):
    rng = np.random.default_rng(2)
    rest = _lattice()
    cs, mem = _clusters(rest, 0.25)
    Q = _rot(rng)
    t = rng.normal(size=3)
    pos = rest.copy()
    pos[:, :3] = rest[:, :3] @ Q.T + t
    vel = np.zeros_like(pos)
    S = _params(alpha=1.0, beta=beta, iters=2)
    p, v = pos.copy(), vel.copy()
    orc.orc_sm_step_d(oracle.ptr(p), oracle.ptr(v), len(p), oracle.ptr(cs), oracle.ptr(mem), len(cs) - 1, oracle.ptr(rest), C.byref(S))
    np.testing.assert_allclose(p[:, :3], pos[:, :3], atol=1e-10)
    np.testing.assert_allclose(v[:, :3], 0.0, atol=1e-6)  # = position residual (1e-10) / dt


def test_alpha_zero_is_pure_prediction(orc):
    rng = np.random.default_rng(3)
    rest = _lattice()
    cs, mem = _clusters(rest, 0.25)
    pos = rest.copy()
    pos[:, :3] += rng.normal(0, 0.01, size=(len(rest), 3))
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.1, size=(len(rest), 3))
    S = _params(alpha=0.0, beta=0.3, iters=2, g=(0, -9.8, 0))
    p, v = pos.copy(), vel.copy()
    orc.orc_sm_step_d(oracle.ptr(p), oracle.ptr(v), len(p), oracle.ptr(cs), oracle.ptr(mem), len(cs) - 1, oracle.ptr(rest), C.byref(S))
    g = np.array([0, -9.8, 0])
    pred = pos[:, :3] + (vel[:, :3] + g * 0.01) * 0.01
    np.testing.assert_allclose(p[:, :3], pred, atol=1e-12)


def test_beta_one_volume_normalised_linear_map(orc):
    """beta = 1: the goal transform is A / sqrt(|det A|) with A = Apq Aqq^-1; a pure scaling diag(s) of the
    rest shape has A = diag(s), so goals are rest scaled by s / sqrt(prod s) about the centre of mass."""
    rest = _lattice(4, 0.1)
    n = len(rest)
    cs = np.array([0, n], np.int32)  # one cluster with everything
    mem = np.arange(n, dtype=np.int32)
    s = np.array([1.2, 0.9, 1.05])
    cm0 = rest[:, :3].mean(0)
    pos = rest.copy()
    pos[:, :3] = (rest[:, :3] - cm0) * s + cm0
    S = _params(alpha=1.0, beta=1.0, iters=1)
    p, v = pos.copy(), np.zeros_like(pos)
    orc.orc_sm_step_d(oracle.ptr(p), oracle.ptr(v), n, oracle.ptr(cs), oracle.ptr(mem), 1, oracle.ptr(rest), C.byref(S))
    expect = (rest[:, :3] - cm0) * (s / np.sqrt(np.prod(s))) + cm0
    np.testing.assert_allclose(p[:, :3], expect, atol=1e-10)


def _masked_step(orc, p, v, cs, mem, rest, S, active):
    p, v = p.astype(np.float32).copy(), v.astype(np.float32).copy()
    rest32 = np.ascontiguousarray(rest, np.float32)
    orc.orc_sm_step_masked_f(oracle.ptr(p), oracle.ptr(v), len(p), oracle.ptr(cs), oracle.ptr(mem), len(cs) - 1,
                             oracle.ptr(rest32), C.byref(S), oracle.ptr(active))
    return p, v


def test_mask_all_active_equals_unmasked(orc):
    rng = np.random.default_rng(5)
    rest = _lattice().astype(np.float32)
    cs, mem = _clusters(rest, 0.25)
    pos = rest.copy()
    pos[:, :3] += rng.normal(0, 0.01, size=(len(rest), 3)).astype(np.float32)
    vel = np.zeros_like(pos)
    S = _params(alpha=0.9, beta=0.3, iters=2, g=(0, -9.8, 0))
    p1, v1 = pos.copy(), vel.copy()
    orc.orc_sm_step_f(oracle.ptr(p1), oracle.ptr(v1), len(p1), oracle.ptr(cs), oracle.ptr(mem), len(cs) - 1, oracle.ptr(rest), C.byref(S))
    p2, v2 = _masked_step(orc, pos, vel, cs, mem, rest, S, np.ones(len(rest), np.int32))
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2)


@pytest.mark.parametrize("beta", [0.0, 0.3])
def test_masked_rigid_motion_is_a_fixed_point(orc, beta):
    # Clusters that lost members must use the rest centre of mass and Aqq of the members left: a rigid motion of
    # the rest shape is then still reproduced exactly by every active particle.  With stale (all-member) rest
    # quantities the goals would be shifted by R (cm0_active - cm0_all).
    rng = np.random.default_rng(6)
    rest = _lattice().astype(np.float32)
    cs, mem = _clusters(rest, 0.25)
    Q = _rot(rng)
    t = rng.normal(size=3)
    pos = rest.copy()
    pos[:, :3] = (rest[:, :3].astype(np.float64) @ Q.T + t).astype(np.float32)
    active = (rng.random(len(rest)) > 0.3).astype(np.int32)
    S = _params(alpha=1.0, beta=beta, iters=2)
    p, v = _masked_step(orc, pos, np.zeros_like(pos), cs, mem, rest, S, active)
    a = active == 1
    np.testing.assert_allclose(p[a, :3], pos[a, :3], atol=2e-5)
    # inactive particles keep their prediction (zero velocity, no gravity here: unchanged)
    np.testing.assert_array_equal(p[~a, :3], pos[~a, :3])


@pytest.mark.parametrize(
    "p0,v0,expect_v",
    [
        # crosses +x: x reflects with restitution 0.9 about the start point, y and z stay at the start point
        ((0.99, 0.5, 0.5), (3.0, 0.2, -0.1), (-0.9 * 3.0, 0.0, 0.0)),
        # crosses the floor: y reflects (gravity included in the reflected step), x and z stay
        ((0.5, 0.01, 0.5), (0.1, -2.0, 0.3), (0.0, -0.9 * (-2.0 - 9.8 * 0.01), 0.0)),
        # crosses x and y (edge): the x branch runs first and resets y to the start point, which is inside,
        # so only x reflects
        ((0.99, 0.01, 0.5), (3.0, -2.0, 0.0), (-0.9 * 3.0, 0.0, 0.0)),
        # stays inside: free flight under gravity
        ((0.5, 0.5, 0.5), (0.1, 0.2, 0.3), (0.1, 0.2 - 9.8 * 0.01, 0.3)),
    ],
)
def test_single_particle_wall_response(orc, p0, v0, expect_v):
    # OrientedPrtObj::UpdateCluster_SM (ElasticObject_OP.cpp:572-608) with a one-member cluster (no matching:
    # ShapeMatching returns early for one vertex).  Box [0, 1]^3, dt = 0.01, g = -9.8 y.
    S = _params(alpha=1.0, beta=0.0, iters=1, g=(0.0, -9.8, 0.0), dt=0.01)
    for k in range(3):
        S.bmin[k], S.bmax[k] = 0.0, 1.0
    p = np.array([[*p0, 0.0]], np.float64)
    v = np.array([[*v0, 0.0]], np.float64)
    rest = p.copy()
    cs, mem = np.array([0, 1], np.int32), np.array([0], np.int32)
    orc.orc_sm_step_d(oracle.ptr(p), oracle.ptr(v), 1, oracle.ptr(cs), oracle.ptr(mem), 1, oracle.ptr(rest), C.byref(S))
    np.testing.assert_allclose(v[0, :3], expect_v, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(p[0, :3], np.array(p0) + 0.01 * np.array(expect_v), rtol=1e-12, atol=1e-14)
    assert np.all((p[0, :3] >= 0.0) & (p[0, :3] <= 1.0))
