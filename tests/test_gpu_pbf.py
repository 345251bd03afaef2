"""PBF entry points vs the oracle of the CPU path (rx_sph_pbd.cpp)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import abi_driver, check, scene
from fltk_sph_ice_amd.abi_driver import GridBuffers
from fltk_sph_ice_amd.device import DevBuf

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _compressed(sc, factor=0.97):
    pos, vel = sc.particles()
    pos[:, :3] *= np.float32(factor)
    pos[:, :3] += np.float32(sc.spacing)
    return pos, vel


def test_lambda_and_dp_parity(lib):
    sc = scene.pbf_dam_break(14, 14, 14)
    pos, vel = _compressed(sc)
    sp = sc.params()
    lib.CuSetParameters(C.byref(sp))
    n = sc.n
    g = GridBuffers(n, sp.NumCells, lib)
    dpos, dvel = DevBuf.from_array(pos), DevBuf.from_array(vel)
    c = g.set_object_to_cell(dpos.ptr, dvel.ptr)
    dd, ds, ddp = DevBuf(4 * n), DevBuf(4 * n), DevBuf(16 * n)
    lib.CuPbdSphScalingFactor(dpos.ptr, dd.ptr, ds.ptr, sc.epsilon, c)
    check(lib)
    P = oracle.params_from(sp, sc)
    orc = oracle.load()
    rd, rl = np.zeros(n, np.float32), np.zeros(n, np.float32)
    orc.orc_pbf_scaling_factor_f(oracle.ptr(pos), n, C.byref(P), oracle.ptr(rd), oracle.ptr(rl))
    np.testing.assert_allclose(dd.download(np.float32), rd, rtol=TOL)
    lam = ds.download(np.float32)
    # lambda = -C/(sd + eps) with C = rho/rho0 - 1 cancelling: per particle (rho_i/rho0 + |C_i|)/(sd_i + eps)
    lscale = oracle.pbf_lambda_scale(pos, P)
    err = oracle.rel_err_vs_scale(lam, rl, lscale)
    print(f"PBF lambda: max per-particle error {err:.2e} of (rho/rho0 + |C|)/(sd + eps)")
    assert err <= TOL, err
    # sign: lambda has the sign of -C
    C_ = rd / sp.Density - 1
    assert np.all(np.sign(rl[np.abs(C_) > 1e-4]) == -np.sign(C_[np.abs(C_) > 1e-4]))
    drl = DevBuf.from_array(rl)  # held until the call returns
    lib.CuPbdSphPositionCorrection(dpos.ptr, drl.ptr, ddp.ptr, c)
    check(lib)
    rdp = np.zeros((n, 4), np.float32)
    orc.orc_pbf_position_correction_f(oracle.ptr(pos), oracle.ptr(rl), n, C.byref(P), np.float32(sp.Dt), oracle.ptr(rdp))
    dp = ddp.download(np.float32).reshape(n, 4)
    dscale = np.maximum(np.linalg.norm(rdp[:, :3], axis=1), oracle.pbf_dp_scale(pos, rl, P, sp.Dt))
    err = oracle.rel_err_vs_scale(dp[:, :3], rdp[:, :3], dscale)
    print(f"PBF dp: max per-particle error {err:.2e} of max(|dp_i|, sum_j |dp_ij|)")
    assert err <= TOL, err


def test_ext_force_and_xsph(lib):
    sc = scene.pbf_dam_break(12, 12, 12)
    pos, vel = sc.particles()
    vel[:, :3] = np.random.default_rng(9).normal(0, 0.3, (sc.n, 3)).astype(np.float32)
    sp = sc.params()
    sp.Viscosity = 0.5
    lib.CuSetParameters(C.byref(sp))
    n = sc.n
    g = GridBuffers(n, sp.NumCells, lib)
    dpos, dvel = DevBuf.from_array(pos), DevBuf.from_array(vel)
    c = g.set_object_to_cell(dpos.ptr, dvel.ptr)
    P = oracle.params_from(sp, sc)
    orc = oracle.load()
    rd = np.zeros(n, np.float32)
    orc.orc_pbf_density_f(oracle.ptr(pos), n, C.byref(P), oracle.ptr(rd))
    dd = DevBuf(4 * n)
    lib.CuPbdSphDensity(dd.ptr, c)
    np.testing.assert_allclose(dd.download(np.float32), rd, rtol=TOL)
    racc = np.zeros((n, 4), np.float32)
    orc.orc_pbf_ext_force_f(oracle.ptr(pos), oracle.ptr(vel), oracle.ptr(rd), n, C.byref(P), oracle.ptr(racc))
    df = DevBuf(16 * n)
    drd = DevBuf.from_array(rd)  # held until the last call that reads it
    lib.CuPbdSphExternalForces(drd.ptr, df.ptr, c, sc.dt)
    check(lib)
    acc = df.download(np.float32).reshape(n, 4)
    ascale = np.maximum(np.linalg.norm(racc[:, :3], axis=1), oracle.pbf_ext_force_scale(pos, vel, rd, P))
    err = oracle.rel_err_vs_scale(acc[:, :3], racc[:, :3], ascale)
    print(f"PBF external force: max per-particle error {err:.2e} of max(|a_i|, mu sum_j |a_ij| + |g|)")
    assert err <= TOL, err
    rv = np.zeros((n, 4), np.float32)
    orc.orc_pbf_xsph_f(oracle.ptr(pos), oracle.ptr(vel), oracle.ptr(rd), n, C.byref(P), np.float32(0.01), oracle.ptr(rv))
    dnv = DevBuf(16 * n)
    lib.CuXSphViscosity(dpos.ptr, dvel.ptr, dnv.ptr, drd.ptr, np.float32(0.01), c)
    check(lib)
    nv = dnv.download(np.float32).reshape(n, 4)
    assert np.abs(nv[:, :3] - rv[:, :3]).max() <= TOL * np.abs(rv[:, :3]).max()
    # in place (dNewVel == dVel): every neighbour velocity is read from the slot-order copy made before any is
    # overwritten, so the result is the out-of-place one, bit for bit
    dvin = DevBuf.from_array(vel)
    lib.CuXSphViscosity(dpos.ptr, dvin.ptr, dvin.ptr, drd.ptr, np.float32(0.01), c)
    check(lib)
    assert np.array_equal(dvin.download(np.float32).reshape(n, 4).view(np.uint32), nv.view(np.uint32))


def test_abi_step_parity(lib):
    sc = scene.pbf_dam_break(14, 14, 14)
    pos, vel = sc.particles()
    P = oracle.params_from(sc.params(), sc)
    rpos, rvel, rdens, st = oracle.pbf_step(pos, vel, P)
    gpos, gvel, gdens, errs = abi_driver.pbf_step(sc, pos, vel)
    scale = np.abs(rpos[:, :3]).max()
    assert np.abs(gpos[:, :3] - rpos[:, :3]).max() <= TOL * scale
    vscale = np.abs(rvel[:, :3]).max()
    assert np.abs(gvel[:, :3] - rvel[:, :3]).max() <= 1e-3 * vscale
    np.testing.assert_allclose(errs, [st.max_err[k] for k in range(sc.iterations)], rtol=1e-3, atol=1e-6)


def test_density_fluctuation_and_stats(lib):
    n = 5000
    rng = np.random.default_rng(1)
    dens = (1000 + rng.normal(0, 20, n)).astype(np.float32)
    dd, de, ds = DevBuf.from_array(dens), DevBuf(4 * n), DevBuf(4 * n)
    m = lib.CuPbdSphCalDensityFluctuation(ds.ptr, de.ptr, dd.ptr, np.float32(1000.0), n)
    check(lib)
    ref = np.maximum(dens.astype(np.float64) - 1000, 0) / 1000
    assert m == pytest.approx(ref.mean(), rel=1e-4)
    mx, mean = C.c_float(), C.c_float()
    lib.CuPbdSphDensityErrorStats(dd.ptr, np.float32(1000.0), n, C.byref(mx), C.byref(mean))
    assert mx.value == pytest.approx(ref.max(), rel=1e-5)
    assert mean.value == pytest.approx(np.abs(dens.astype(np.float64) - 1000).mean() / 1000, rel=1e-4)
