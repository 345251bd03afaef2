"""Fused WCSPH solver (rxSPH_GPU, graph replay) vs the oracle of the CPU path (rx_sph_cpu.cpp)."""
import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import _abi, scene
from fltk_sph_ice_amd.solvers import SPHSolver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", [False, True])
def test_steps_parity(graph):
    sc = scene.wcsph_dam_break(16, 16, 16)
    sc.viscosity = 3.5 / 998.29  # stable multi-step setting (see test_gpu_sph.py)
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    rp, rv = pos, vel
    for _ in range(5):
        rp, rv, rd = oracle.sph_step(rp, rv, P, sc.dt)
    s = SPHSolver(sc, pos, vel, graph=graph)
    s.step(5)
    gp, gv, gd = s.array(_abi.RX_ARRAY_POSITION), s.array(_abi.RX_ARRAY_VELOCITY), s.array(_abi.RX_ARRAY_DENSITY)
    np.testing.assert_allclose(gd, rd, rtol=1e-5)
    assert np.abs(gp[:, :3] - rp[:, :3]).max() <= 1e-5
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-3 * np.abs(rv[:, :3]).max()


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 1)])
def test_tiny_fluids_parity(dims):
    sc = scene.wcsph_dam_break(*dims)
    sc.viscosity = 3.5 / 998.29
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    rp, rv = pos, vel
    for _ in range(3):
        rp, rv, rd = oracle.sph_step(rp, rv, P, sc.dt)
    s = SPHSolver(sc, pos, vel, graph=False)
    s.step(3)
    gp, gv, gd = s.array(_abi.RX_ARRAY_POSITION), s.array(_abi.RX_ARRAY_VELOCITY), s.array(_abi.RX_ARRAY_DENSITY)
    np.testing.assert_allclose(gd, rd, rtol=1e-5)
    assert np.abs(gp[:, :3] - rp[:, :3]).max() <= 1e-5
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-3 * np.abs(rv[:, :3]).max()


def test_graph_and_eager_agree_bitwise():
    sc = scene.wcsph_dam_break(20, 12, 9)
    pos, vel = sc.particles()
    a = SPHSolver(sc, pos, vel, graph=True)
    b = SPHSolver(sc, pos, vel, graph=False)
    for _ in range(3):  # several replays with eager work and syncs between them
        a.step(4)
        b.step(4)
        a.sync()
        b.sync()
        assert np.array_equal(a.array(_abi.RX_ARRAY_POSITION), b.array(_abi.RX_ARRAY_POSITION))


def test_drift_against_fp64_report():
    sc = scene.wcsph_dam_break(12, 12, 12)
    sc.viscosity = 3.5 / 998.29  # stable multi-step setting (see test_gpu_sph.py)
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    p64, v64 = pos.astype(np.float64), vel.astype(np.float64)
    steps = 200
    for _ in range(steps):
        p64, v64, _ = oracle.sph_step(p64, v64, P, sc.dt, double=True)
    s = SPHSolver(sc, pos, vel)
    s.step(steps)
    gp = s.array(_abi.RX_ARRAY_POSITION)
    drift = np.abs(gp[:, :3] - p64[:, :3]).max() / sc.h
    print(f"WCSPH {steps}-step FP32 GPU vs FP64 oracle max drift: {drift:.3e} h")
    assert np.isfinite(drift)


def test_wcsph_at_32768_particles():
    """The WCSPH configuration at its full size: a 32^3 block (32,768 particles) at spacing h/2 in the unit box,
    h = 0.02, k = 3, dt = 1e-3, with the stable viscosity 3.5/rho0 of the multi-step tests.  One step matches the
    oracle (density to 1e-5 relative, positions to 1e-5); over 200 steps the state stays finite and inside the
    box.  The fastest speed and the density range are reported: with k = 3 the fluid is soft, and the collapse
    drives pressure waves whose speeds exceed the block's free-fall speed."""
    sc = scene.wcsph_dam_break(32, 32, 32)
    assert sc.n == 32768
    sc.viscosity = 3.5 / 998.29
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    rp, rv, rd = oracle.sph_step(pos, vel, P, sc.dt)
    s = SPHSolver(sc, pos, vel, graph=True)
    s.step(1)
    gp, gd = s.array(_abi.RX_ARRAY_POSITION), s.array(_abi.RX_ARRAY_DENSITY)
    np.testing.assert_allclose(gd, rd, rtol=1e-5)
    assert np.abs(gp[:, :3] - rp[:, :3]).max() <= 1e-5
    s.step(199)
    gp, gv, gd = s.array(_abi.RX_ARRAY_POSITION), s.array(_abi.RX_ARRAY_VELOCITY), s.array(_abi.RX_ARRAY_DENSITY)
    s.close()
    vmax = float(np.abs(gv[:, :3]).max())
    print(f"WCSPH, 32,768 particles, 200 steps: max |v| {vmax:.3f}, density {gd.min():.1f} .. {gd.max():.1f}")
    assert np.isfinite(gp).all() and np.isfinite(gv).all() and np.isfinite(gd).all()
    lo, hi = np.array(sc.box_min, np.float32), np.array(sc.box_max, np.float32)
    assert (gp[:, :3] >= lo - 1e-5).all() and (gp[:, :3] <= hi + 1e-5).all()
