"""Pins the oracle's legacy ice launches (reference slot layout): the iteration pass (UpdateCluster_SM_Itr) keeps
rigidly moved shapes, undoes reflections of clusters of at least 10 members and clamps to the walls; the first pass
equals the batched oracle step on disjoint clusters; the per-particle average returns a common position and
leaves particles without clusters alone."""
import numpy as np
import pytest

import oracle
from oracle import legacy_ice as LI


def _rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def _lattice(n=4, s=0.1, origin=(0.0, 0.0, 0.0)):
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3) * s + np.asarray(origin)
    p = np.zeros((len(g), 4))
    p[:, :3] = g
    return p


def _one_cluster(rest):
    n = len(rest)
    return np.array([0, n], np.int32), np.arange(n, dtype=np.int32)


def _itr(L, sld, S, double=True):
    T = np.float64 if double else np.float32
    cur = np.zeros(3 * L["nslots"], T)
    ccm = np.zeros(3 * L["nc"], T)
    dummy4 = np.zeros((L["n"], 4), T)
    vel = np.zeros(3 * L["nslots"], T)
    LI.cluster_pass(L, dummy4, dummy4, np.ascontiguousarray(sld, T), cur, ccm, vel, 0.01, S, itr=True, double=double)
    return cur.reshape(-1, 3), ccm, vel


@pytest.mark.parametrize("n", [2, 4])
def test_iteration_keeps_rigid_motion(n):
    rng = np.random.default_rng(1)
    rest = _lattice(n)
    cs, mem = _one_cluster(rest)
    L = LI.layout(rest, cs, mem)
    Q, t = _rot(rng), rng.normal(0, 0.5, 3)
    sld = rest[:, :3] @ Q.T + t
    cur, ccm, vel = _itr(L, sld.ravel(), LI.sm_params(0.01))
    np.testing.assert_allclose(cur, sld, atol=1e-8)  # Jacobi stops at an off-diagonal sum of 1e-8
    np.testing.assert_allclose(ccm, sld.mean(0), atol=1e-12)
    assert not vel.any()  # the iteration pass leaves the velocities alone


@pytest.mark.parametrize("n,undone", [(4, True), (2, False)])
def test_iteration_reflection_fix_needs_ten_members(n, undone):
    # A cluster mirrored through its centre plane z: det Apq < 0.  With at least 10 members (4^3 = 64) the third
    # column of Apq is negated and the goal is the unmirrored rest shape about the current centre; with 8 members
    # (2^3) the polar factor is the reflection itself and the mirrored shape is kept.
    rest = _lattice(n)
    cs, mem = _one_cluster(rest)
    L = LI.layout(rest, cs, mem)
    c = rest[:, :3].mean(0)
    sld = rest[:, :3].copy()
    sld[:, 2] = 2 * c[2] - sld[:, 2]
    sld[:, 2] += 0.01 * (np.arange(len(sld)) % 3 == 0)  # break the symmetry that would make the image coincide
    cur, _, _ = _itr(L, sld.ravel(), LI.sm_params(0.01))
    cm = sld.mean(0)
    unmirrored = rest[:, :3] - c + cm
    if undone:
        assert np.abs(cur - unmirrored).max() < 0.02
        assert np.abs(cur - sld).max() > 0.05
    else:
        assert np.abs(cur - sld).max() < 0.02
        assert np.abs(cur - unmirrored).max() > 0.05


def test_iteration_clamps_goals_to_the_walls():
    rest = _lattice(3)
    cs, mem = _one_cluster(rest)
    L = LI.layout(rest, cs, mem)
    sld = rest[:, :3] + np.array([0.0, -1.05, 0.0])  # the block pushed 0.05 below a floor at y = -1
    S = LI.sm_params(0.01, lo=(-5.0, -1.0, -5.0), hi=(5.0, 5.0, 5.0))
    cur, _, _ = _itr(L, sld.ravel(), S)
    expect = sld.copy()
    expect[:, 1] = np.maximum(expect[:, 1], -1.0)  # rigid goal = sld, then clamped
    np.testing.assert_allclose(cur, expect, atol=1e-8)


def test_first_pass_equals_batched_step_on_disjoint_clusters(orc):
    # UpdateCluster_SM in the slot layout equals the batched oracle step with alpha = 1, beta = 0 and one
    # iteration when no particle is shared (the average is then the identity)
    import ctypes as C

    rng = np.random.default_rng(4)
    nc, k = 12, 7
    rest = np.zeros((nc * k, 4))
    for c in range(nc):
        rest[c * k:(c + 1) * k, :3] = rng.normal(0, 0.1, (k, 3)) + c * 0.7 - 4.0
    pos = rest.copy()
    pos[:, :3] += rng.normal(0, 0.01, pos[:, :3].shape)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.5, (len(pos), 3))
    cs = np.arange(0, nc * k + 1, k, dtype=np.int32)
    mem = np.arange(nc * k, dtype=np.int32)
    L = LI.layout(rest, cs, mem)
    S = LI.sm_params(0.01, lo=(-4.2, -5.0, -5.0), hi=(5.0, 5.0, 5.0))  # some particles reach the wall at x = -4.2
    cur = np.zeros(3 * L["nslots"])
    ccm = np.zeros(3 * nc)
    v3 = np.zeros(3 * L["nslots"])
    LI.cluster_pass(L, pos, vel, np.zeros(3 * len(pos)), cur, ccm, v3, 0.01, S, itr=False, double=True)
    p, v = pos.copy(), vel.copy()
    orc.orc_sm_step_d(oracle.ptr(p), oracle.ptr(v), len(p), oracle.ptr(cs), oracle.ptr(mem), nc, oracle.ptr(rest), C.byref(S))
    np.testing.assert_allclose(cur.reshape(-1, 3), p[:, :3], atol=1e-12)
    np.testing.assert_allclose(v3.reshape(-1, 3), v[:, :3], atol=1e-9)


def test_average_of_common_positions_and_empty_particles():
    rng = np.random.default_rng(5)
    rest = _lattice(3)
    n = len(rest)
    # clusters: every particle with its x-neighbours; particle 0 is left out of every cluster
    cs, mem = [0], []
    for p in range(1, n):
        m = [q for q in range(1, n) if abs(rest[q, 0] - rest[p, 0]) < 0.15 and np.allclose(rest[q, 1:3], rest[p, 1:3])]
        mem += m
        cs.append(len(mem))
    cs, mem = np.array(cs, np.int32), np.array(mem, np.int32)
    L = LI.layout(rest, cs, mem)
    target = rng.normal(0, 1, (n, 3)).astype(np.float32)
    tvel = rng.normal(0, 1, (n, 3)).astype(np.float32)
    sm_pos = target[L["pidx"]].ravel().copy()  # every cluster agrees on every member's position
    sm_vel = tvel[L["pidx"]].ravel().copy()
    sld = np.full(3 * n, 7.0, np.float32)
    sldv = np.full(3 * n, -7.0, np.float32)
    LI.average(L, sld, sldv, sm_pos, sm_vel)
    np.testing.assert_allclose(sld.reshape(n, 3)[1:], target[1:], rtol=1e-6)
    np.testing.assert_allclose(sldv.reshape(n, 3)[1:], tvel[1:], rtol=1e-6)
    assert np.all(sld[:3] == 7.0) and np.all(sldv[:3] == -7.0)  # in no cluster: unchanged
    # distinct member positions: the mean over the particle's memberships (and -1 entries are skipped)
    sm_pos = rng.normal(0, 1, 3 * L["nslots"]).astype(np.float32)
    ptoc = L["ptoc"].reshape(n, L["ptoc_param_size"], L["ptoc_max"])
    L2 = dict(L)
    p2 = ptoc.copy()
    p2[5, 0, 0] = -1  # drop the first membership of particle 5
    L2["ptoc"] = p2.ravel().copy()
    LI.average(L2, sld, sldv, sm_pos, sm_vel)
    slots = [L["indx"][2 * c] + o for c, o in zip(ptoc[5, 0, 1:L["ptoc_indx"][5]], ptoc[5, 1, 1:L["ptoc_indx"][5]])]
    np.testing.assert_allclose(sld.reshape(n, 3)[5], sm_pos.reshape(-1, 3)[slots].astype(np.float64).mean(0), rtol=1e-6)
