"""Marching cubes through the C ABI (CuMCCalTriNum / CuMCCalEdgeVrts / CuMCCalTri / CuMCCalNrm, in the order of
rxMCMeshGPU::CreateMeshV) against the CPU path: every stage array bit-exact, vertices and normals included."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import check, scene
from fltk_sph_ice_amd.abi_driver import GridBuffers
from fltk_sph_ice_amd.device import DevBuf
from fltk_sph_ice_amd.mesh import MCMesh

pytestmark = pytest.mark.gpu


def _field(n, h, lo, f):
    z, y, x = np.meshgrid(*[lo[a] + np.arange(n[a], dtype=np.float64) * h[a] for a in (2, 1, 0)], indexing="ij")
    return f(x, y, z).astype(np.float32)


def _compare(m: MCMesh, vol, thr):
    print(f"mc lattice {m.n} threshold {thr}", flush=True)
    nv, nt = m.create_mesh(vol, thr)
    r = oracle.mc_mesh(vol, m.n, m.h, m.lo, thr)
    assert (m.active, nv, nt) == (r["active"], r["nvrts"], r["ntris"])
    g = m.stage_arrays()
    for k in ("voxbit", "vnum", "tnum", "occ", "occscan", "eocc", "escan"):
        assert np.array_equal(g[k], r[k]), k
    if r["active"]:
        # the vertex and triangle scans are only computed when some voxel is occupied
        for k in ("vscan", "tscan"):
            assert np.array_equal(g[k], r[k]), k
        assert np.array_equal(g["compacted"], r["compacted"][: r["active"]])
    occ = r["eocc"].astype(bool)
    assert np.array_equal(g["edge_vrts"][occ], r["edge_vrts"][occ])
    assert np.array_equal(m.vertices(), r["vrts"])
    assert np.array_equal(m.triangles(), r["tris"])
    assert np.array_equal(m.normals(), r["nrms"])
    return r


@pytest.mark.parametrize("shape", ["sphere", "torus", "noise", "plane"])
def test_stages_match_cpu_path(lib, shape):
    n, h, lo = (37, 29, 23), (0.07, 0.08, 0.09), (-1.3, -1.1, -1.0)
    if shape == "sphere":
        vol, thr = _field(n, h, lo, lambda x, y, z: np.sqrt(x * x + y * y + z * z)), 0.713
    elif shape == "torus":
        vol, thr = _field(n, h, lo, lambda x, y, z: np.sqrt((np.sqrt(x * x + y * y) - 0.61) ** 2 + z * z)), 0.297
    elif shape == "noise":
        vol, thr = np.random.default_rng(7).random(int(np.prod(n)), dtype=np.float32), 0.5
    else:
        vol, thr = _field(n, h, lo, lambda x, y, z: 0.7 * x - 0.4 * y + 0.5 * z), 0.1
    m = MCMesh(n, lo, h, lib)
    r = _compare(m, vol, thr)
    assert r["ntris"] > 100
    # the same object again with another level (stale arrays from the first call must not leak in)
    _compare(m, vol, thr * 0.9 if shape != "noise" else 0.3)


def test_small_and_empty_lattices(lib):
    for n in [(1, 1, 1), (2, 2, 2), (3, 1, 2), (5, 4, 3)]:
        m = MCMesh(n, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), lib)
        k = int(np.prod(n))
        rng = np.random.default_rng(k)
        _compare(m, rng.random(k, dtype=np.float32), 0.5)
        for v in (0.0, 2.0):  # nothing below / everything below: no surface
            _compare(m, np.full(k, v, np.float32), 1.0)
        nv, nt = m.create_mesh(np.full(k, 1.0, np.float32), 1.0)  # level equal to every sample
        assert (m.active, nv, nt) == (0, 0, 0)


def test_surface_of_a_particle_density_lattice(lib):
    # the reference's use: a density lattice of the fluid (CuSphGridDensity), then the mesh at a level
    sc = scene.pbf_dam_break(16, 16, 16)
    pos, vel = sc.particles()
    sp = sc.params()
    lib.CuSetParameters(C.byref(sp))
    g = GridBuffers(sc.n, sp.NumCells, lib)
    dpos, dvel = DevBuf.from_array(pos), DevBuf.from_array(vel)
    c = g.set_object_to_cell(dpos.ptr, dvel.ptr)
    n, d = (40, 32, 36), np.float32(0.03)
    lo = (np.float32(-0.1), np.float32(-0.08), np.float32(-0.09))
    vol = DevBuf(4 * int(np.prod(n)))
    lib.CuSphGridDensity(vol.ptr, c, n[0], n[1], n[2], lo[0], lo[1], lo[2], d, d, d)
    check(lib)
    m = MCMesh(n, lo, (float(d),) * 3, lib)
    thr = 0.5 * float(sp.Density)
    m.create_mesh(vol, thr)
    r = _compare(m, vol.download(np.float32), thr)
    assert r["ntris"] > 1000


def test_argument_errors_are_reported(lib):
    from fltk_sph_ice_amd._abi import float3, uint3

    lib.CuClearLastError()
    a, v, t = C.c_uint(), C.c_uint(), C.c_uint()
    buf = DevBuf(64)
    # num_voxels that does not match the lattice
    lib.CuMCCalTriNum(buf.ptr, buf.ptr, buf.ptr, buf.ptr, buf.ptr, buf.ptr, buf.ptr, buf.ptr, buf.ptr, uint3(2, 2, 2), 9,
                      float3(1, 1, 1), 0.5, C.byref(a), C.byref(v), C.byref(t))
    assert lib.CuGetLastErrorCode() == 2 and (a.value, v.value, t.value) == (0, 0, 0)
    lib.CuClearLastError()
