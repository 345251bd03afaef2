"""Neighbour/surface detection and LaunchHeatTransferGPU on the GPU vs the oracle (bit-exact)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import check, scene
from fltk_sph_ice_amd.abi_driver import GridBuffers
from fltk_sph_ice_amd.device import DevBuf

pytestmark = pytest.mark.gpu


def _detect(lib, sc, pos):
    sp = sc.params()
    lib.CuSetParameters(C.byref(sp))
    n = len(pos)
    g = GridBuffers(n, sp.NumCells, lib)
    dpos = DevBuf.from_array(pos)
    dvel = DevBuf.from_array(np.zeros_like(pos))
    c = g.set_object_to_cell(dpos.ptr, dvel.ptr)
    dnb = DevBuf(4 * 30 * n)
    lib.CuSphDetectNeighborParticles(dnb.ptr, c, np.float32(sc.h), n)
    check(lib)
    return sp, c, dpos, dnb


def test_detect_neighbours_and_surface(lib):
    sc = scene.pbf_dam_break(14, 12, 10, seed=3)
    pos, _ = sc.particles()
    pos[::17, :3] += np.float32(0.3 * sc.spacing)  # denser spots: some lists reach 30
    n = len(pos)
    sp, c, dpos, dnb = _detect(lib, sc, pos)
    nb = dnb.download(np.int32).reshape(n, 30)
    P = oracle.params_from(sp, sc)
    orc = oracle.load()
    ref = np.zeros((n, 30), np.int32)
    orc.orc_detect_neighbors_f(oracle.ptr(pos), n, C.byref(P), oracle.ptr(ref))
    assert np.array_equal(nb, ref)
    assert (ref[:, 29] == 30).any() and (ref[:, 29] < 30).any()
    # surface flags, twice (the second call starts from the first call's flags)
    dsurf = DevBuf.from_array(np.zeros(n, np.int32))
    rsurf = np.zeros(n, np.int32)
    for _ in range(2):
        lib.CuSphDetectSurfaceParticles(dnb.ptr, dsurf.ptr, c, dpos.ptr, np.float32(sc.spacing))
        check(lib)
        orc.orc_detect_surface_f(oracle.ptr(ref), oracle.ptr(rsurf), oracle.ptr(pos), n, np.float32(sp.Mass),
                                 np.float32(sc.spacing))
        assert np.array_equal(dsurf.download(np.int32), rsurf)
    assert (rsurf > 0).any() and (rsurf == 0).any()


@pytest.mark.parametrize("seed", [0, 1])
def test_heat_transfer_bit_exact(lib, seed):
    sc = scene.pbf_dam_break(16, 12, 12, seed=5 + seed)
    pos, _ = sc.particles()
    n = len(pos)
    sp, c, dpos, dnb = _detect(lib, sc, pos)
    nb = dnb.download(np.int32).reshape(n, 30)
    rng = np.random.default_rng(seed)
    temps = rng.uniform(230.0, 270.0, n).astype(np.float32)
    heats = rng.uniform(-5.0, 180.0, n).astype(np.float32)
    phase = rng.choice(np.array([-2, -1, 1, 2], np.int32), n)
    surf = rng.integers(-1, 25, n).astype(np.int32)
    dens = rng.uniform(0.0, 1200.0, n).astype(np.float32)
    dens[::13] = 0.01  # below the 0.05 floor
    melt = -np.ones(1000, np.int32)
    melt[0] = 40
    melt[1:41] = rng.integers(0, n, 40)
    changed = np.zeros(n, np.int32)
    D = {k: DevBuf.from_array(v) for k, v in dict(heats=heats, temps=temps, dtemps=np.zeros(n, np.float32), surf=surf,
                                                  dens=dens, phase=phase, changed=changed, melt=melt).items()}
    args = (np.float32(0.9), np.float32(20.0), np.float32(0.2), np.float32(0.35), np.float32(1.0 / 60.0))
    for step in range(3):
        lib.LaunchHeatTransferGPU(D["heats"].ptr, D["temps"].ptr, D["dtemps"].ptr, D["surf"].ptr, dnb.ptr, D["dens"].ptr,
                                  D["phase"].ptr, D["changed"].ptr, D["melt"].ptr, *args, n)
        check(lib)
    orc = oracle.load()
    rh, rt, rd, rp, rc = heats.copy(), temps.copy(), np.zeros(n, np.float32), phase.copy(), changed.copy()
    for step in range(3):
        orc.orc_heat_transfer_gpu_f(oracle.ptr(rh), oracle.ptr(rt), oracle.ptr(rd), oracle.ptr(surf), oracle.ptr(nb),
                                    oracle.ptr(dens), oracle.ptr(rp), oracle.ptr(rc), oracle.ptr(melt), *args, n)
    assert np.array_equal(D["phase"].download(np.int32), rp)
    assert np.array_equal(D["changed"].download(np.int32), rc)
    np.testing.assert_array_equal(D["dtemps"].download(np.float32), rd)
    np.testing.assert_array_equal(D["temps"].download(np.float32), rt)
    np.testing.assert_array_equal(D["heats"].download(np.float32), rh)
    assert len(set(rp.tolist())) >= 3 and rc.sum() > 0
