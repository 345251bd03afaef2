"""Anisotropic kernels through the C ABI (CuSphCalUpdatedPosition, CuSphCalCovarianceMatrix, CuSphSVDecomposition,
CuSphCalTransformMatrix, CuSphGridDensityAniso) against the CPU path on the same sorted grid: bit-exact.  The
surface test follows the reference's sequence (rxSPH_GPU::CalAnisotropicKernel, rx_sph_gpu.cpp:2036-2095, then
the lattice of rx_sph_gpu.cpp:2940-2965)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import _abi, check, scene
from fltk_sph_ice_amd.abi_driver import GridBuffers
from fltk_sph_ice_amd.device import DevBuf
from fltk_sph_ice_amd.mesh import MCMesh

pytestmark = pytest.mark.gpu


def _setup(lib, nx, ny, nz, jitter, seed, outside=0.0):
    sc = scene.pbf_dam_break(nx, ny, nz)
    pos, vel = sc.particles()
    rng = np.random.default_rng(seed)
    pos[:, :3] += rng.uniform(-jitter, jitter, (sc.n, 3)).astype(np.float32) * np.float32(sc.spacing)
    if outside:
        # a share of the particles moved past the grid's low faces: the clamped hash puts them in face cells
        k = rng.random(sc.n) < outside
        pos[k, :3] -= rng.uniform(0.05, 0.2, (int(k.sum()), 3)).astype(np.float32)
    sp = sc.params()
    lib.CuSetParameters(C.byref(sp))
    g = GridBuffers(sc.n, sp.NumCells, lib)
    dpos, dvel = DevBuf.from_array(pos), DevBuf.from_array(vel)
    c = g.set_object_to_cell(dpos.ptr, dvel.ptr)
    stage = (g.spos.download(np.float32).reshape(-1, 4), g.index.download(np.uint32), g.cstart.download(np.uint32),
             g.cend.download(np.uint32))
    geo = np.array([*sp.WorldOrigin, *sp.CellWidth], np.float32)
    gsize = np.array([*sp.GridSize], np.int32)
    return sc, sp, c, (g, dpos, dvel), stage, geo, gsize


@pytest.mark.parametrize("jitter,outside", [(0.0, 0.0), (0.3, 0.0), (0.3, 0.05)])
def test_stages_match_cpu_path(lib, jitter, outside):
    sc, sp, c, keep, (spos, sidx, cs, ce), geo, gsize = _setup(lib, 14, 12, 10, jitter, 3, outside)
    n = sc.n
    h, lam = 2.0 * float(sp.EffectiveRadius), 0.9  # rxSPH_GPU::CalAnisotropicKernel: h = 2 x effective radius
    print(f"aniso n={n} jitter={jitter}", flush=True)
    up, wpos, posw = DevBuf(16 * n), DevBuf(16 * n), DevBuf(16 * n)
    cmat, eig, rmat, gmat = DevBuf(36 * n), DevBuf(12 * n), DevBuf(36 * n), DevBuf(36 * n)
    lib.CuSphCalUpdatedPosition(up.ptr, wpos.ptr, lam, h, c)
    lib.CuSphCalCovarianceMatrix(posw.ptr, cmat.ptr, h, c)
    lib.CuSphSVDecomposition(cmat.ptr, posw.ptr, eig.ptr, rmat.ptr, n)
    lib.CuSphCalTransformMatrix(eig.ptr, rmat.ptr, gmat.ptr, n)
    check(lib)
    nlat = (36, 30, 28)
    lo = (np.float32(-0.02), np.float32(-0.02), np.float32(-0.02))
    d = np.float32(0.5 * sp.EffectiveRadius)
    grid = DevBuf(4 * int(np.prod(nlat)))
    lib.CuSphGridDensityAniso(grid.ptr, gmat.ptr, 0.0, c, *nlat, *lo, d, d, d)
    check(lib)
    r = oracle.aniso_pipeline(spos, sidx, cs, ce, geo, gsize, lam, h, float(sp.EffectiveRadius), float(sp.Mass), float(sp.Wpoly6),
                              (nlat, lo, (d, d, d)))
    got = {"up": up.download(np.float32).reshape(n, 4), "wpos": wpos.download(np.float32).reshape(n, 4),
           "posw": posw.download(np.float32).reshape(n, 4), "cmat": cmat.download(np.float32).reshape(n, 9),
           "eigen": eig.download(np.float32).reshape(n, 3), "rmat": rmat.download(np.float32).reshape(n, 9),
           "gmat": gmat.download(np.float32).reshape(n, 9), "grid": grid.download(np.float32)}
    for k, v in got.items():
        assert np.array_equal(v, r[k]), (k, np.abs(v - r[k]).max())
    # the neighbourhoods are full (more than 10 neighbours inside the block) and the lattice sees the fluid
    assert np.median(r["posw"][:, 3]) > 10
    assert (r["grid"] > 0).mean() > 0.03


def test_anisotropic_surface_mesh(lib):
    # the reference's use: anisotropic lattice density, then marching cubes at a level
    sc, sp, c, (g, dpos, dvel), stage, geo, gsize = _setup(lib, 16, 16, 12, 0.2, 8)
    n = sc.n
    h = 2.0 * float(sp.EffectiveRadius)
    up, wpos, posw = DevBuf(16 * n), DevBuf(16 * n), DevBuf(16 * n)
    cmat, eig, rmat, gmat = DevBuf(36 * n), DevBuf(12 * n), DevBuf(36 * n), DevBuf(36 * n)
    lib.CuSphCalUpdatedPosition(up.ptr, wpos.ptr, 0.9, h, c)
    c = g.set_object_to_cell(up.ptr, dvel.ptr)  # the covariance is taken over the grid of the smoothed centres
    lib.CuSphCalCovarianceMatrix(posw.ptr, cmat.ptr, h, c)
    lib.CuSphSVDecomposition(cmat.ptr, posw.ptr, eig.ptr, rmat.ptr, n)
    lib.CuSphCalTransformMatrix(eig.ptr, rmat.ptr, gmat.ptr, n)
    nlat, d = (40, 40, 32), np.float32(0.5 * sp.EffectiveRadius)
    lo = (np.float32(-0.03),) * 3
    c = g.set_object_to_cell(dpos.ptr, dvel.ptr)  # the lattice is summed over the particle positions
    vol = DevBuf(4 * int(np.prod(nlat)))
    lib.CuSphGridDensityAniso(vol.ptr, gmat.ptr, 0.0, c, *nlat, *lo, d, d, d)
    check(lib)
    m = MCMesh(nlat, lo, (float(d),) * 3, lib)
    nv, nt = m.create_mesh(vol, 0.5 * float(sp.Density))
    assert nt > 1000
    V = m.vertices()[:, :3]
    block_hi = np.array(sc.particles()[0][:, :3].max(0))
    # the surface wraps the block: vertices within about a particle spacing of it
    assert V.min() > -0.03 and np.all(V.max(0) < block_hi + 3 * sc.spacing)


def test_lattice_of_no_particles_is_zero(lib):
    # an empty fluid: the cell arrays are unset and every lattice point reads 0
    sp = scene.pbf_dam_break(4, 4, 4).params()
    lib.CuSetParameters(C.byref(sp))
    c = _abi.rxParticleCell()
    c.uNumCells = sp.NumCells
    nlat = (9, 7, 5)
    vol = DevBuf.from_array(np.full(int(np.prod(nlat)), 7.0, np.float32))
    d = np.float32(0.05)
    lib.CuSphGridDensityAniso(vol.ptr, None, 0.0, c, *nlat, 0.0, 0.0, 0.0, d, d, d)
    check(lib)
    assert np.array_equal(vol.download(np.float32), np.zeros(int(np.prod(nlat)), np.float32))
