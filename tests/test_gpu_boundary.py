"""Boundary-particle entry points (rx_cu_funcs.cuh:81-83, 131-135) vs the oracle: a fluid block in the corner of
its box, boundary particles on the six box faces, grids built through the ABI as rxSPH_GPU::InitBoundary does."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import abi_driver, check, scene
from fltk_sph_ice_amd.abi_driver import BoundaryGrid, GridBuffers
from fltk_sph_ice_amd.device import DevBuf

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _setup(lib, dims=(12, 12, 12), factor=0.97):
    sc = scene.pbf_dam_break(*dims)
    pos, vel = sc.particles()
    pos[:, :3] *= np.float32(factor)  # slightly compressed, half a spacing off the walls
    pos[:, :3] += np.float32(0.5 * sc.spacing)
    sp = sc.params()
    abi_driver.set_boundary_grid(sp, sc.box_min, sc.box_max, sc.h)
    lib.CuSetParameters(C.byref(sp))
    b = scene.box_surface_particles(sc.box_min, sc.box_max, 0.5 * sc.h)
    bg = BoundaryGrid(sp, b, lib)
    g = GridBuffers(sc.n, sp.NumCells, lib)
    dpos, dvel = DevBuf.from_array(pos), DevBuf.from_array(vel)
    c = g.set_object_to_cell(dpos.ptr, dvel.ptr)
    P = oracle.params_from(sp, sc)
    rvol = np.zeros(len(b), np.float32)
    oracle.load().orc_boundary_volume_f(oracle.ptr(b), len(b), C.byref(P), oracle.ptr(rvol))
    return sc, sp, pos, vel, b, bg, g, c, dpos, dvel, P, rvol


def test_boundary_grid_and_volumes(lib):
    sc, sp, pos, vel, b, bg, g, c, dpos, dvel, P, rvol = _setup(lib)
    vol = bg.vol.download(np.float32)
    np.testing.assert_allclose(vol, rvol, rtol=TOL)
    # the sorted boundary grid holds every boundary particle once, in cell order
    idx = bg.g.index.download(np.uint32)
    assert np.array_equal(np.sort(idx), np.arange(len(b), dtype=np.uint32))
    hs = bg.g.hash.download(np.uint32)
    assert np.all(np.diff(hs.astype(np.int64)) >= 0) and hs.max() < sp.NumCellsB
    # corner and edge particles have fewer boundary neighbours than face particles, hence larger volumes
    assert vol.max() > 1.2 * np.median(vol)


def test_wcsph_density_and_forces_with_boundary(lib):
    sc, sp, pos, vel, b, bg, g, c, dpos, dvel, P, rvol = _setup(lib)
    n, orc = sc.n, oracle.load()
    dd, dp = DevBuf(4 * n), DevBuf(4 * n)
    lib.CuSphDensity(dd.ptr, dp.ptr, c)
    lib.CuSphBoundaryDensity(dd.ptr, dp.ptr, dpos.ptr, bg.vol.ptr, bg.cell, n)
    check(lib)
    rd, rp = np.zeros(n, np.float32), np.zeros(n, np.float32)
    orc.orc_sph_density_f(oracle.ptr(pos), n, C.byref(P), oracle.ptr(rd), oracle.ptr(rp))
    fluid_only = rd.copy()
    orc.orc_boundary_density_f(oracle.ptr(pos), n, oracle.ptr(b), oracle.ptr(rvol), len(b), C.byref(P), oracle.ptr(rd),
                               oracle.ptr(rp))
    assert (rd > fluid_only).sum() > 0.2 * n  # the walls reach a good part of the block
    np.testing.assert_allclose(dd.download(np.float32), rd, rtol=TOL)
    assert np.abs(dp.download(np.float32) - rp).max() <= TOL * sp.GasStiffness * rd.max()
    # forces: fluid pairs, then the boundary pressure term added on top
    vel[:, :3] = np.random.default_rng(2).normal(0, 0.1, (n, 3)).astype(np.float32)
    racc = np.zeros((n, 4), np.float32)
    orc.orc_sph_force_f(oracle.ptr(pos), oracle.ptr(vel), oracle.ptr(rd), oracle.ptr(rp), n, C.byref(P), oracle.ptr(racc))
    only = racc.copy()
    orc.orc_boundary_force_f(oracle.ptr(pos), n, oracle.ptr(b), oracle.ptr(rvol), len(b), C.byref(P), oracle.ptr(rd),
                             oracle.ptr(rp), oracle.ptr(racc))
    assert np.abs(racc - only).max() > 0
    dvel2 = DevBuf.from_array(vel)
    c = g.set_object_to_cell(dpos.ptr, dvel2.ptr)
    drd, drp, df = DevBuf.from_array(rd), DevBuf.from_array(rp), DevBuf(16 * n)
    lib.CuSphForces(drd.ptr, drp.ptr, df.ptr, c, sc.dt)
    lib.CuSphBoundaryForces(drd.ptr, drp.ptr, dpos.ptr, bg.vol.ptr, df.ptr, bg.cell, n)
    check(lib)
    acc = df.download(np.float32).reshape(n, 4)
    fscale = oracle.boundary_force_scale(pos, b, rvol, rd, rp, P, base=oracle.sph_force_scale(pos, vel, rd, rp, P))
    scale = np.maximum(np.linalg.norm(racc[:, :3], axis=1), fscale)
    err = oracle.rel_err_vs_scale(acc[:, :3], racc[:, :3], scale)
    print(f"WCSPH force with boundary: max per-particle error {err:.2e}")
    assert err <= TOL, err


def test_pbf_boundary_density(lib):
    sc, sp, pos, vel, b, bg, g, c, dpos, dvel, P, rvol = _setup(lib)
    n, orc = sc.n, oracle.load()
    dd = DevBuf(4 * n)
    lib.CuPbdSphDensity(dd.ptr, c)
    lib.CuPbdSphBoundaryDensity(dd.ptr, dpos.ptr, bg.vol.ptr, bg.cell, n)
    check(lib)
    rd = np.zeros(n, np.float32)
    orc.orc_pbf_density_f(oracle.ptr(pos), n, C.byref(P), oracle.ptr(rd))
    orc.orc_boundary_density_f(oracle.ptr(pos), n, oracle.ptr(b), oracle.ptr(rvol), len(b), C.byref(P), oracle.ptr(rd), None)
    np.testing.assert_allclose(dd.download(np.float32), rd, rtol=TOL)


def test_pbf_multipliers_and_correction_with_boundary(lib):
    sc, sp, pos, vel, b, bg, g, c, dpos, dvel, P, rvol = _setup(lib)
    n, nb, orc = sc.n, len(b), oracle.load()
    dd, ds = DevBuf(4 * n), DevBuf(4 * n)
    lib.CuPbdSphScalingFactorWithBoundary(dpos.ptr, dd.ptr, ds.ptr, np.float32(sc.epsilon), c, bg.vol.ptr, bg.scl.ptr, bg.cell)
    check(lib)
    rd, rl, rlb = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(nb, np.float32)
    orc.orc_pbf_scaling_factor_wb_f(oracle.ptr(pos), n, oracle.ptr(b), oracle.ptr(rvol), nb, C.byref(P), oracle.ptr(rd),
                                    oracle.ptr(rl), oracle.ptr(rlb))
    np.testing.assert_allclose(dd.download(np.float32), rd, rtol=TOL)
    lam, lamb = ds.download(np.float32), bg.scl.download(np.float32)
    # lambda = -C/(sd + eps) with C = rho/rho0 - 1 cancelling: per particle (rho_i/rho0 + |C_i|)/(sd_i + eps)
    ls, lbs = oracle.pbf_lambda_scale(pos, P, b, rvol)
    el, elb = oracle.rel_err_vs_scale(lam, rl, ls), oracle.rel_err_vs_scale(lamb, rlb, lbs)
    print(f"PBF multipliers with boundary: max per-particle error fluid {el:.2e}, boundary {elb:.2e}")
    assert el <= TOL and elb <= TOL, (el, elb)
    assert np.isfinite(lamb).all()
    # the multipliers differ from the fluid-only ones next to the walls
    d0, l0 = np.zeros(n, np.float32), np.zeros(n, np.float32)
    orc.orc_pbf_scaling_factor_f(oracle.ptr(pos), n, C.byref(P), oracle.ptr(d0), oracle.ptr(l0))
    assert (np.abs(l0 - rl) > 1e-3 * np.abs(rl).max()).sum() > 0.1 * n
    # position correction from the oracle's multipliers
    ddp = DevBuf(16 * n)
    drl, drlb = DevBuf.from_array(rl), DevBuf.from_array(rlb)  # held until the call returns
    lib.CuPbdSphPositionCorrectionWithBoundary(dpos.ptr, drl.ptr, ddp.ptr, c, bg.vol.ptr, drlb.ptr, bg.cell)
    check(lib)
    rdp = np.zeros((n, 4), np.float32)
    orc.orc_pbf_position_correction_wb_f(oracle.ptr(pos), oracle.ptr(rl), n, oracle.ptr(b), oracle.ptr(rvol), oracle.ptr(rlb),
                                         nb, C.byref(P), np.float32(sp.Dt), oracle.ptr(rdp))
    dp = ddp.download(np.float32).reshape(n, 4)
    dscale = np.maximum(np.linalg.norm(rdp[:, :3], axis=1), oracle.pbf_dp_scale(pos, rl, P, sp.Dt, b, rvol, rlb))
    err = oracle.rel_err_vs_scale(dp[:, :3], rdp[:, :3], dscale)
    print(f"PBF correction with boundary: max per-particle error {err:.2e}")
    assert err <= TOL, err


def test_mismatched_boundary_grid_is_an_error(lib):
    sc, sp, pos, vel, b, bg, g, c, dpos, dvel, P, rvol = _setup(lib, dims=(6, 6, 6))
    n = sc.n
    bad = abi_driver.rxParticleCell.from_buffer_copy(bg.cell)
    bad.uNumCells = bg.cell.uNumCells - 1
    dens = DevBuf.from_array(np.ones(n, np.float32))
    lib.CuClearLastError()
    lib.CuPbdSphBoundaryDensity(dens.ptr, dpos.ptr, bg.vol.ptr, bad, n)
    assert lib.CuGetLastErrorCode() != 0
    lib.CuClearLastError()
    assert np.all(dens.download(np.float32) == 1.0)
