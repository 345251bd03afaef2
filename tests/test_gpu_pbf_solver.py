"""Fused PBF solver (rxPBDSPH_GPU) vs the oracle of the CPU path: same step, same inputs."""
import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import _abi, scene
from fltk_sph_ice_amd.solvers import PBFSolver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", [(14, 14, 14), (20, 9, 11)])
def test_one_step_parity(dims):
    sc = scene.pbf_dam_break(*dims)
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    rpos, rvel, rdens, st = oracle.pbf_step(pos, vel, P)
    s = PBFSolver(sc, pos, vel)
    s.step(1)
    gpos, gvel = s.positions(), s.velocities()
    scale = np.abs(rpos[:, :3]).max()
    assert np.abs(gpos[:, :3] - rpos[:, :3]).max() <= 1e-5 * scale
    assert np.abs(gvel[:, :3] - rvel[:, :3]).max() <= 1e-3 * np.abs(rvel[:, :3]).max()
    stt = s.stats()
    for k in range(sc.iterations):
        assert stt[1 + k] == pytest.approx(st.max_err[k], rel=1e-3, abs=1e-6)
    assert stt[18] == 0  # no list overflow


def test_multi_step_parity_and_list_reuse():
    """Ten steps: the list built on x* is reused by the next step's predict pass."""
    sc = scene.pbf_dam_break(16, 16, 16)
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    rp, rv = pos, vel
    for _ in range(10):
        rp, rv, _, _ = oracle.pbf_step(rp, rv, P)
    s = PBFSolver(sc, pos, vel)
    s.step(10)
    gp = s.positions()
    err = np.abs(gp[:, :3] - rp[:, :3]).max() / sc.h
    print(f"10-step FP32 GPU vs FP32 oracle: {err:.2e} h")
    assert err < 1e-3
    assert s.stats()[22] == 0


def test_lambda_sum_recorded():
    sc = scene.pbf_dam_break(12, 12, 12)
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    _, _, _, _, lam = oracle.pbf_step(pos, vel, P, want_lambda=True)
    s = PBFSolver(sc, pos, vel)
    s.option(_abi.RX_PBF_OPT_RECORD_LAMBDA, 1)
    s.step(1)
    g = s.array(_abi.RX_ARRAY_LAMBDA_SUM)
    assert np.abs(g - lam).max() <= 1e-3 * np.abs(lam).max()
