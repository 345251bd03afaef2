"""Batched shape matching (rxIceSM) and the legacy launch function vs the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import load, scene
from fltk_sph_ice_amd.device import DevBuf
from fltk_sph_ice_amd.solvers import IceSM

pytestmark = pytest.mark.gpu


def _make(sc, graph=True):
    pos, vel = sc.particles()
    sm = IceSM(pos, alpha=sc.alpha, beta=sc.beta, dt=sc.dt, gravity=sc.gravity, box_min=sc.box_min, box_max=sc.box_max,
               iterations=sc.iterations, cluster_radius=sc.cluster_radius, graph=graph)
    sm.set_state(pos, vel)
    return sm, pos, vel


def _oracle_steps(sc, pos, vel, cs, mem, steps):
    S = oracle.OrcSmParams()
    S.alpha, S.beta, S.dt = sc.alpha, sc.beta, sc.dt
    for k in range(3):
        S.gravity[k], S.bmin[k], S.bmax[k] = sc.gravity[k], sc.box_min[k], sc.box_max[k]
    S.iterations, S.volume_conservation, S.restitution = sc.iterations, 1, 0.9
    p, v = pos.copy(), vel.copy()
    rest = pos.copy()
    for _ in range(steps):
        oracle.load().orc_sm_step_f(oracle.ptr(p), oracle.ptr(v), len(p), oracle.ptr(cs), oracle.ptr(mem), len(cs) - 1,
                                    oracle.ptr(rest), C.byref(S))
    return p, v


def test_clusters_are_radius_neighbourhoods():
    sc = scene.ice_block(10, 9, 7)
    sm, pos, _ = _make(sc)
    cs, mem, hist = sm.clusters()
    d = np.linalg.norm(pos[:, None, :3].astype(np.float64) - pos[None, :, :3].astype(np.float64), axis=-1)
    for i in range(sc.n):
        ref = np.nonzero(d[i] <= sc.cluster_radius)[0]
        assert np.array_equal(mem[cs[i]:cs[i + 1]], ref)
    sizes = np.diff(cs)
    print("cluster sizes: min", sizes.min(), "max", sizes.max())
    assert hist.sum() == sc.n


@pytest.mark.parametrize("steps", [1, 5])
def test_batched_step_parity(steps):
    sc = scene.ice_block(12, 10, 8)
    sm, pos, vel = _make(sc)
    cs, mem, _ = sm.clusters()
    rp, rv = _oracle_steps(sc, pos, vel, cs, mem, steps)
    sm.step(steps)
    gp, gv = sm.state()
    assert np.abs(gp[:, :3] - rp[:, :3]).max() <= 2e-5 * np.abs(rp[:, :3]).max()
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 2e-3 * max(np.abs(rv[:, :3]).max(), 1e-3)


def test_autotuned_tiles_and_graph_give_same_result():
    sc = scene.ice_block(14, 12, 10)
    a, pos, vel = _make(sc, graph=True)
    b, _, _ = _make(sc, graph=False)
    tiles = a.autotune()
    assert set(tiles.tolist()) <= {8, 16, 32, 64}
    b.set_tiles([64, 64, 64, 64, 64])
    a.set_state(pos, vel)
    for _ in range(3):
        a.step(2)
        b.step(2)
        a.sync()
        b.sync()
    pa, _ = a.state()
    pb, _ = b.state()
    assert np.abs(pa - pb).max() <= 1e-5 * np.abs(pb).max()


def test_rigid_rotation_preserved():
    sc = scene.ice_block(8, 8, 8)
    sc.gravity = (0.0, 0.0, 0.0)
    sc.alpha, sc.beta = 1.0, 0.3
    sm, pos, _ = _make(sc)
    th = 0.3
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]], np.float64)
    c = pos[:, :3].astype(np.float64).mean(0)
    rot = pos.copy()
    rot[:, :3] = ((pos[:, :3] - c) @ Q.T + c).astype(np.float32)
    sm.set_state(rot, np.zeros_like(rot))
    sm.step(1)
    gp, gv = sm.state()
    assert np.abs(gp[:, :3] - rot[:, :3]).max() < 2e-5


def test_legacy_launch_shape_matching_disjoint_clusters(lib):
    """LaunchShapeMatchingGPU (Ice_SM.h:23-36 layout) on disjoint clusters equals the oracle step with
    alpha = 1, beta = 0 and one iteration (averaging is the identity for disjoint clusters)."""
    sc = scene.wcsph_dam_break(4, 4, 4)
    sp = sc.params()
    sp.BoundaryMin[0] = sp.BoundaryMin[1] = sp.BoundaryMin[2] = -5.0
    sp.BoundaryMax[0] = sp.BoundaryMax[1] = sp.BoundaryMax[2] = 5.0
    sp.Gravity[1] = -9.8
    lib.CuSetParameters(C.byref(sp))
    rng = np.random.default_rng(7)
    nc, k = 40, 9
    n = nc * k
    rest = np.zeros((n, 4), np.float32)
    for c in range(nc):
        rest[c * k:(c + 1) * k, :3] = rng.normal(0, 0.1, (k, 3)) + c * 1.0 - 2.0
    pos = rest.copy()
    pos[:, :3] += rng.normal(0, 0.01, (n, 3))
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.2, (n, 3))
    dt = np.float32(0.01)
    cs = np.arange(0, n + 1, k, dtype=np.int32)
    mem = np.arange(n, dtype=np.int32)
    indx = np.stack([cs[:-1], cs[1:] - 1], 1).astype(np.int32).ravel()
    org = rest[:, :3].astype(np.float32).ravel().copy()
    ocm = np.stack([rest[c * k:(c + 1) * k, :3].astype(np.float64).mean(0) for c in range(nc)]).astype(np.float32).ravel()
    bufs = {name: DevBuf.from_array(a) for name, a in
            dict(pos=pos, vel=vel, org=org, cur=np.zeros(3 * n, np.float32), ocm=ocm, ccm=np.zeros(3 * nc, np.float32),
                 v=np.zeros(3 * n, np.float32), pidx=mem, indx=indx).items()}
    lib.LaunchShapeMatchingGPU(nc, bufs["pos"].ptr, bufs["vel"].ptr, bufs["org"].ptr, bufs["cur"].ptr, bufs["ocm"].ptr,
                               bufs["ccm"].ptr, bufs["v"].ptr, bufs["pidx"].ptr, bufs["indx"].ptr, dt)
    cur = bufs["cur"].download(np.float32).reshape(n, 3)
    S = oracle.OrcSmParams()
    S.alpha, S.beta, S.dt = 1.0, 0.0, float(dt)
    for a in range(3):
        S.gravity[a], S.bmin[a], S.bmax[a] = (0, -9.8, 0)[a], -5.0, 5.0
    S.iterations, S.volume_conservation, S.restitution = 1, 1, 0.9
    p, v = pos.copy(), vel.copy()
    oracle.load().orc_sm_step_f(oracle.ptr(p), oracle.ptr(v), n, oracle.ptr(cs), oracle.ptr(mem), nc, oracle.ptr(rest), C.byref(S))
    np.testing.assert_allclose(cur, p[:, :3], atol=2e-6)
    vv = bufs["v"].download(np.float32).reshape(n, 3)
    np.testing.assert_allclose(vv, v[:, :3], atol=5e-4)


@pytest.mark.parametrize("graph,radius", [(False, 2.5), (True, 2.5), (False, 1.6), (True, 1.6)])
def test_masked_step_parity(graph, radius):
    """Melt mask: about a third of the particles are no longer ice; clusters use their ice members only.
    Radius 2.5 spacings: 3 steps.  Radius 1.6 spacings: clusters of 7-19 lose a third of their members, and a
    partial cluster of 4-6 nearly coplanar members has cond(Aqq) up to 6e5; its linear term (beta = 0.3)
    multiplies the FP32 rounding of the positions by about that much in each later step (the oracle works in
    FP64 between FP32 inputs and outputs), so the bar is checked over one step there."""
    sc = scene.ice_block(12, 10, 8)
    sc.radius_spacings = radius
    steps = 3 if radius > 2.0 else 1
    sm, pos, vel = _make(sc, graph=graph)
    cs, mem, _ = sm.clusters()
    active = (np.random.default_rng(11).random(sc.n) > 0.35).astype(np.int32)
    sm.set_active(active)
    S = oracle.OrcSmParams()
    S.alpha, S.beta, S.dt = sc.alpha, sc.beta, sc.dt
    for k in range(3):
        S.gravity[k], S.bmin[k], S.bmax[k] = sc.gravity[k], sc.box_min[k], sc.box_max[k]
    S.iterations, S.volume_conservation, S.restitution = sc.iterations, 1, 0.9
    rp, rv, rest = pos.copy(), vel.copy(), pos.copy()
    for _ in range(steps):
        oracle.load().orc_sm_step_masked_f(oracle.ptr(rp), oracle.ptr(rv), len(rp), oracle.ptr(cs), oracle.ptr(mem), len(cs) - 1,
                                           oracle.ptr(rest), C.byref(S), oracle.ptr(active))
    sm.step(steps)
    gp, gv = sm.state()
    ptol = 2e-5 * np.abs(rp[:, :3]).max()
    assert np.abs(gp[:, :3] - rp[:, :3]).max() <= ptol
    # v = (x_new - x_old) / dt: at radius 1.6 the bar is the position bar over dt
    vtol = 2e-3 * max(np.abs(rv[:, :3]).max(), 1e-3) if radius > 2.0 else ptol / sc.dt
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= vtol
    sm.set_active(None)  # mask removed: back to the unmasked step
    sm.set_state(pos, vel)
    sm.step(1)
    up, _ = sm.state()
    rp2, rv2 = _oracle_steps(sc, pos, vel, cs, mem, 1)
    assert np.abs(up[:, :3] - rp2[:, :3]).max() <= 2e-5 * np.abs(rp2[:, :3]).max()


def test_graded_block_fills_every_size_bin():
    """A block at spacing s beside one at spacing s/2, clusters within 1.6 s: 7 to 19 members in the coarse part,
    up to about 140 in the dense part, and every size between at the seam, so all five size bins of the batched
    launch (<= 8, 16, 32, 64 and more) hold clusters.  Autotuned tiles and the graph step, against the oracle."""
    sc = scene.ice_block(10, 8, 8)
    s = sc.spacing
    rng = np.random.default_rng(21)

    def lattice(n3, sp, origin):
        g = np.stack(np.meshgrid(*[np.arange(k) for k in n3], indexing="ij"), -1).reshape(-1, 3)
        return np.asarray(origin) + (g + 0.5) * sp + rng.uniform(-0.01 * sp, 0.01 * sp, g.shape)

    coarse = lattice((10, 8, 8), s, sc.origin)
    dense = lattice((14, 16, 16), 0.5 * s, (sc.origin[0] + 10 * s, sc.origin[1], sc.origin[2]))
    p3 = np.concatenate([coarse, dense]).astype(np.float32)
    n = len(p3)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = p3
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0.0, 0.1, (n, 3)).astype(np.float32)
    assert (p3 > np.asarray(sc.box_min) + s).all() and (p3 < np.asarray(sc.box_max) - s).all()
    sm = IceSM(pos, alpha=sc.alpha, beta=sc.beta, dt=sc.dt, gravity=sc.gravity, box_min=sc.box_min, box_max=sc.box_max,
               iterations=sc.iterations, cluster_radius=sc.cluster_radius, graph=True)
    sm.set_state(pos, vel)
    cs, mem, _ = sm.clusters()
    sizes = np.diff(cs)
    edges = [0, 8, 16, 32, 64, 1 << 30]
    per_bin = [int(((sizes > edges[b]) & (sizes <= edges[b + 1])).sum()) for b in range(5)]
    print("cluster sizes", sizes.min(), "-", sizes.max(), "per launch bin", per_bin)
    assert all(c > 0 for c in per_bin)
    tiles = sm.autotune()
    sm.set_state(pos, vel)
    rp, rv = _oracle_steps(sc, pos, vel, cs, mem, 3)
    sm.step(3)
    gp, gv = sm.state()
    print("tiles", tiles.tolist())
    assert np.abs(gp[:, :3] - rp[:, :3]).max() <= 2e-5 * np.abs(rp[:, :3]).max()
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 2e-3 * max(np.abs(rv[:, :3]).max(), 1e-3)
    sm.close()


def test_ice_block_at_131072_particles():
    """The batched shape-matching configuration at its full size: 131,072 particles, clusters within 1.6
    spacings, alpha = 0.9, beta = 0.3, 2 iterations.  One step matches the oracle to the parity bar of the
    smaller blocks.  Over 50 more steps (the block falls 0.5 onto the floor): the state stays finite, the block's
    centre falls, every cluster keeps its members, the fastest particle stays below three times the free-fall
    speed of the block's top, and no particle sinks a lattice spacing into a wall (the wall rule acts on the
    prediction; the shape-matching goal can carry a particle slightly past it, as in the reference)."""
    sc = scene.ice_block()
    assert sc.n == 131072
    sm, pos, vel = _make(sc)
    cs, mem, hist = sm.clusters()
    rp, rv = _oracle_steps(sc, pos, vel, cs, mem, 1)
    sm.step(1)
    gp, gv = sm.state()
    ep = np.abs(gp[:, :3] - rp[:, :3]).max() / np.abs(rp[:, :3]).max()
    print(f"ice block, 131,072 particles: one-step position error {ep:.2e} of the box scale; cluster sizes "
          f"{ {i: int(c) for i, c in enumerate(hist) if c} }")
    assert ep <= 2e-5
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 2e-3 * max(np.abs(rv[:, :3]).max(), 1e-3)
    sm.step(50)
    gp, gv = sm.state()
    cs2, mem2, _ = sm.clusters()
    lo, hi = np.array(sc.box_min, np.float32), np.array(sc.box_max, np.float32)
    sink = float(max((lo - gp[:, :3]).max(), (gp[:, :3] - hi).max(), 0.0))
    top = float(pos[:, 1].max() - lo[1])
    vmax = float(np.abs(gv[:, :3]).max())
    print(f"after 51 steps: centre height {gp[:, 1].mean():.3f} (from {pos[:, 1].mean():.3f}), deepest past a wall "
          f"{sink:.4f}, max |v| {vmax:.2f}")
    assert np.isfinite(gp).all() and np.isfinite(gv).all()
    assert gp[:, 1].mean() < pos[:, 1].mean()
    assert np.array_equal(cs, cs2) and np.array_equal(mem, mem2)
    assert vmax < 3.0 * np.sqrt(2.0 * 9.8 * top)
    assert sink < sc.spacing
