"""1000-step drift of the FP32 GPU solvers against the FP64 oracle of the same step (reported, not gated):
WCSPH at the dam-break parameters on a reduced block, and PBF through the oracle's FP64 step.  The one-step
parity tests bound the per-step difference; over many steps the two trajectories separate at the rate the
flow amplifies rounding, which these tests print at 10, 100 and 1000 steps (in units of h)."""
import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import _abi, scene
from fltk_sph_ice_amd.solvers import PBFSolver, SPHSolver

pytestmark = pytest.mark.gpu
MARKS = (10, 100, 1000)


def _report(name, drifts):
    print(f"{name}: FP32 GPU vs FP64 oracle, max position drift / h: " + ", ".join(f"{k} steps {v:.3e}" for k, v in drifts.items()))


def test_wcsph_1000_step_drift_report():
    # the dam-break parameters (h = 0.02, k = 3, dt = 1e-3, spacing h/2, jitter 0.01 h) on a 12^3 block; the
    # viscosity is 3.5 / rho0 because 3.5 itself is beyond the explicit viscous limit of the reference's term
    sc = scene.wcsph_dam_break(12, 12, 12)
    sc.viscosity = 3.5 / 998.29
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    p64, v64 = pos.astype(np.float64), vel.astype(np.float64)
    s = SPHSolver(sc, pos, vel)
    drifts, done = {}, 0
    for mark in MARKS:
        for _ in range(mark - done):
            p64, v64, _ = oracle.sph_step(p64, v64, P, sc.dt, double=True)
        s.step(mark - done)
        done = mark
        gp = s.array(_abi.RX_ARRAY_POSITION)
        assert np.isfinite(gp).all()
        drifts[mark] = float(np.abs(gp[:, :3] - p64[:, :3]).max() / sc.h)
    s.close()
    _report("WCSPH 1728 particles", drifts)
    assert all(np.isfinite(v) for v in drifts.values())


@pytest.mark.parametrize("grad_mass", [0, 1])
def test_pbf_1000_step_drift_report(grad_mass):
    # the PBF parameters (h = 0.1, eps = 600, s_corr, XSPH c = 0.01, dt = 1/120, 4 iterations) on an 8^3 block
    sc = scene.pbf_dam_break(8, 8, 8)
    sc.grad_mass = grad_mass
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    p64, v64 = pos.astype(np.float64), vel.astype(np.float64)
    s = PBFSolver(sc, pos, vel)
    drifts, done = {}, 0
    for mark in MARKS:
        for _ in range(mark - done):
            p64, v64, _, _ = oracle.pbf_step(p64, v64, P, double=True)
        s.step(mark - done)
        done = mark
        gp = s.array(_abi.RX_ARRAY_POSITION)
        assert np.isfinite(gp).all()
        drifts[mark] = float(np.abs(gp[:, :3] - p64[:, :3]).max() / sc.h)
    assert s.stats()[53] == 0
    s.close()
    _report(f"PBF 512 particles ({'mass-carrying' if grad_mass else 'reference'} gradient)", drifts)
    assert all(np.isfinite(v) for v in drifts.values())
