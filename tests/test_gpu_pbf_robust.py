"""Ownership and failure reporting of the fused PBF solver (rxPBDSPH_GPU):

* the learned warm start keeps its own copy of the network: changing or destroying the network after
  rxPBDSPH_GPU_SetNetwork leaves the solver's steps bit-identical to a solver whose network is untouched;
* a step whose particles move beyond the reach of the capped fallback walk is never silent: the step's flag
  and the run flag are set, the next Update returns error 3 once the host has seen it, and the solver
  continues from a fresh grid.
"""
import ctypes as C

import numpy as np
import pytest

from fltk_sph_ice_amd import _abi, learned, scene
from fltk_sph_ice_amd.solvers import PBFSolver

pytestmark = pytest.mark.gpu


def _net(seed):
    n = learned.Network(seed=seed)
    n.set_normalization(np.zeros(12, np.float32), np.ones(12, np.float32), 1.0, 0.0)
    return n


def _learned(sc, net):
    return learned.learned_solver(sc, net, max_iters=2, tol=0.0, warm_scale=0.05)


def test_set_network_keeps_its_own_copy():
    sc = scene.pbf_dam_break(16, 16, 16)
    keep, changed, gone = _net(3), _net(3), _net(3)
    s_keep, s_changed, s_gone = _learned(sc, keep), _learned(sc, changed), _learned(sc, gone)
    changed.set_params(np.zeros(changed.np, np.float32))  # after SetNetwork: must not reach the solver
    changed.set_normalization(np.ones(12, np.float32), np.full(12, 3.0, np.float32), 7.0, 1.0)
    gone.close()
    churn = [_net(s) for s in (11, 12, 13)]  # reuse the freed device memory with other weights
    for s in (s_keep, s_changed, s_gone):
        s.step(3)
    p = [s.positions() for s in (s_keep, s_changed, s_gone)]
    assert np.isfinite(p[0]).all()
    assert np.array_equal(p[0], p[1]) and np.array_equal(p[0], p[2])
    # the copy is a real copy of the weights: a solver given the zeroed network differs
    s_zero = _learned(sc, changed)
    s_zero.step(3)
    assert not np.array_equal(s_zero.positions(), p[0])
    for s in (s_keep, s_changed, s_gone, s_zero):
        s.close()
    for n in churn + [keep, changed]:
        n.close()


def test_capped_walk_is_reported_and_recovers(lib):
    # 216 particles squeezed into a cube of 0.006 (1/8 of the spacing of their lattice) in the middle of a
    # 4 x 2 x 2 box, with the mass-carrying gradient and a small relaxation: the first corrections throw them
    # apart by more than the 4-cell reach of the fallback walk (0.46 at cell 1.15 h).
    sc = scene.pbf_dam_break(6, 6, 6, box=(4.0, 2.0, 2.0))
    sc.epsilon = 1.0
    sc.grad_mass = 1
    pos, vel = sc.particles()
    c = pos[:, :3].mean(0)
    pos[:, :3] = (c + (pos[:, :3] - c) * np.float32(0.02) + np.float32([2.0, 1.0, 1.0]) - c).astype(np.float32)
    s = PBFSolver(sc, pos, vel)
    s.step(1)
    s.sync()
    st = s.stats()
    assert st[22] == 1 and st[53] == 1, (st[22], st[53])
    lib.CuClearLastError()
    # step s applies the records of steps up to s - 2: the call for step 2 reports the capped step 0 (and runs
    # nothing); the call for step 1 runs normally
    rcs = []
    for _ in range(2):
        rcs.append(s.lib.rxPBDSPH_GPU_Update(s.h, C.c_float(sc.dt), 1))
        if rcs[-1] == 3:
            break
    assert rcs[-1] == 3, rcs
    assert b"reach" in lib.CuGetLastErrorString()
    lib.CuClearLastError()
    # the solver carries on from a fresh grid (a step that again outruns the reach is reported by the call
    # after it, which does not step); the run flag stays set until a reset
    codes = []
    for _ in range(4):
        codes.append(s.lib.rxPBDSPH_GPU_Update(s.h, C.c_float(sc.dt), 1))
        s.sync()
        lib.CuClearLastError()
    assert set(codes) <= {0, 3} and 0 in codes, codes
    assert np.isfinite(s.positions()).all()
    assert s.stats()[53] == 1
    s.option(_abi.RX_PBF_OPT_RESET_RUN, 1)
    assert s.stats()[53] == 0
    s.close()
