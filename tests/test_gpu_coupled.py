"""Fluid + melting ice (rxCoupled): one coupled step is the thermal PBF step of every particle followed by the
masked shape-matching substep of the still-frozen ice particles, checked piece by piece."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import _abi, scene
from fltk_sph_ice_amd.solvers import Coupled, IceSM

pytestmark = pytest.mark.gpu


def _small():
    return scene.coupled_melt((24, 16, 16), (12, 16, 12), seed=4)


def test_first_step_is_fluid_step_then_masked_shape_matching():
    cs = _small()
    pos, vel, temps, phase = cs.particles()
    nf, ni = cs.n_fluid, cs.n_ice
    c = Coupled(cs, (pos, vel, temps, phase))
    c.step(1)
    cp, cv = c.array(_abi.RX_ARRAY_POSITION), c.array(_abi.RX_ARRAY_VELOCITY)
    cph = c.array(_abi.RX_ARRAY_PHASE)
    ice_env = c.ice_env
    c.close()

    # the fluid step alone, same inputs
    fl = cs.fluid
    env = fl.env()
    env.max_particles = cs.n
    lib = _abi.load()
    h = lib.rxPBDSPH_GPU_Create(C.byref(env))
    assert h
    assert lib.rxPBDSPH_GPU_SetParticles(h, pos.ctypes.data, vel.ctypes.data, cs.n) == 0
    # the coupled object's list settings (same grid, same summation order)
    assert lib.rxPBDSPH_GPU_SetOption(h, _abi.RX_PBF_OPT_MAX_NEIGHBORS, np.float32(128)) == 0
    assert lib.rxPBDSPH_GPU_SetOption(h, _abi.RX_PBF_OPT_SKIN, np.float32(0.25)) == 0
    assert lib.rxPBDSPH_GPU_SetThermal(h, temps.ctypes.data, None, phase.ctypes.data) == 0
    assert lib.rxPBDSPH_GPU_SetOption(h, _abi.RX_PBF_OPT_HEAT_TD, np.float32(cs.heat_td)) == 0
    assert lib.rxPBDSPH_GPU_Update(h, np.float32(fl.dt), 1) == 0
    fp = np.zeros((cs.n, 4), np.float32)
    fv = np.zeros((cs.n, 4), np.float32)
    fph = np.zeros(cs.n, np.float32)
    assert lib.rxPBDSPH_GPU_GetArray(h, _abi.RX_ARRAY_POSITION, fp.ctypes.data, cs.n) == 0
    assert lib.rxPBDSPH_GPU_GetArray(h, _abi.RX_ARRAY_VELOCITY, fv.ctypes.data, cs.n) == 0
    assert lib.rxPBDSPH_GPU_GetArray(h, _abi.RX_ARRAY_PHASE, fph.ctypes.data, cs.n) == 0
    lib.rxPBDSPH_GPU_Destroy(h)
    assert np.array_equal(cph, fph)
    assert np.array_equal(cp[:nf], fp[:nf]) and np.array_equal(cv[:nf], fv[:nf])

    # the masked shape-matching substep of the ice particles from the post-fluid state
    rest = np.ascontiguousarray(pos[nf:])
    sm = IceSM(rest, alpha=cs.alpha, beta=cs.beta, dt=fl.dt, gravity=fl.gravity, box_min=tuple(ice_env.box_min),
               box_max=tuple(ice_env.box_max), iterations=cs.sm_iterations, cluster_radius=ice_env.cluster_radius)
    cst, mem, _ = sm.clusters()
    sm.close()
    active = (fph[nf:] < 0).astype(np.int32)
    S = oracle.OrcSmParams()
    S.alpha, S.beta, S.dt = cs.alpha, cs.beta, fl.dt
    for k in range(3):
        S.gravity[k], S.bmin[k], S.bmax[k] = fl.gravity[k], ice_env.box_min[k], ice_env.box_max[k]
    S.iterations, S.volume_conservation, S.restitution = cs.sm_iterations, 1, 0.9
    rp, rv = fp[nf:].copy(), fv[nf:].copy()
    oracle.load().orc_sm_step_masked_f(oracle.ptr(rp), oracle.ptr(rv), ni, oracle.ptr(cst), oracle.ptr(mem), len(cst) - 1,
                                       oracle.ptr(rest), C.byref(S), oracle.ptr(active))
    a = active == 1
    assert a.sum() > 0.9 * ni
    scale = np.abs(rp[:, :3]).max()
    assert np.abs(cp[nf:][a, :3] - rp[a, :3]).max() <= 2e-5 * scale
    assert np.abs(cv[nf:][a, :3] - rv[a, :3]).max() <= 2e-3 * max(np.abs(rv[a, :3]).max(), 1e-3)
    # particles no longer frozen keep the fluid result
    assert np.array_equal(cp[nf:][~a], fp[nf:][~a])


def test_coupled_run_melts_ice_at_the_contact():
    cs = _small()
    sp = cs.fluid.spacing
    # ice 1.5 spacings from the water (within h): heat crosses the interface from the first step
    cs.ice_origin = (cs.fluid.block[0] * sp + 1.5 * sp, cs.ice_origin[1], cs.ice_origin[2])
    c = Coupled(cs)
    c.autotune()
    c.step(40)
    T = c.array(_abi.RX_ARRAY_TEMPERATURE)
    ph = c.array(_abi.RX_ARRAY_PHASE)
    p = c.array(_abi.RX_ARRAY_POSITION)
    st = c.stats()
    c.close()
    assert np.isfinite(p).all() and np.isfinite(T).all()
    assert T.min() >= 0.0 and T.max() <= 500.0
    ice = ph[cs.n_fluid:]
    print("ice phases after 40 steps:", {int(k): int((ice == k).sum()) for k in (-2, -1, 1, 2)}, "bad flag", st[22])
    assert (ice != -2).sum() > 0  # heat from the water reached the ice
    assert (ice == -2).sum() > 0.5 * cs.n_ice  # most of the block is still frozen
    assert st[22] == 0  # the neighbour lists stayed within reach of the wide stencil
