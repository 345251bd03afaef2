"""Polygon-obstacle integrators (CuSphIntegrateWithPolygon / CuPbdSphIntegrateWithPolygon) and lattice
densities (CuSphGridDensity / CuPbdSphGridDensity / CuIceMeshMake) vs the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import check, scene
from fltk_sph_ice_amd.abi_driver import GridBuffers, PolygonCells
from fltk_sph_ice_amd.device import DevBuf

pytestmark = pytest.mark.gpu


def _tilted_square(L=0.6):
    """A square of side L split along its diagonal, tilted about two axes, centred in the unit box."""
    nrm = np.array([0.3, 1.0, -0.2])
    nrm /= np.linalg.norm(nrm)
    e1 = np.cross(nrm, [0.0, 0.0, 1.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(nrm, e1)  # e1 x e2 = nrm: the triangles below face +nrm
    o = np.array([0.5, 0.5, 0.5]) - 0.5 * L * (e1 + e2)
    vrts = np.array([o, o + L * e1, o + L * (e1 + e2), o + L * e2], np.float32)
    tris = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    assert np.dot(np.cross(vrts[1] - vrts[0], vrts[2] - vrts[0]), nrm) > 0
    return vrts, tris, o, e1, e2, nrm, L


def _particles(n, seed, dt):
    """Particles in front of, behind and far from the square; crossing points kept 0.02 L away from its edges and
    its diagonal so that the float and double hit decisions agree."""
    vrts, tris, o, e1, e2, nrm, L = _tilted_square()
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(0.03, 0.97, n), rng.uniform(0.03, 0.97, n)
    keep = np.abs(a - b) > 0.02
    a, b = a[keep], b[keep]
    m = len(a)
    speed = rng.uniform(1.0, 3.0, m)
    depth = rng.uniform(0.1, 0.9, m) * speed * dt       # crossing within the step
    kind = rng.integers(0, 3, m)                           # 0: crossing, 1: from behind, 2: not reaching
    side = np.where(kind == 1, -1.0, 1.0)
    depth = np.where(kind == 2, speed * dt * 1.5, depth)
    x = o + (a * L)[:, None] * e1 + (b * L)[:, None] * e2 + (side * depth)[:, None] * nrm
    v = (-side * speed)[:, None] * nrm
    pos = np.zeros((m, 4), np.float32)
    vel = np.zeros((m, 4), np.float32)
    pos[:, :3], vel[:, :3] = x, v
    pos[:, 3] = 1.0
    vel[:, 3] = 2.0
    return pos, vel, vrts, tris, kind


def _params(lib, e):
    sc = scene.wcsph_dam_break(8, 8, 8)
    sp = sc.params()
    sp.Restitution = e
    lib.CuSetParameters(C.byref(sp))
    return sc, sp


@pytest.mark.parametrize("e", [0.0, 0.4])
def test_sph_integrate_with_polygon(lib, e):
    dt = np.float32(0.01)
    pos, vel, vrts, tris, kind = _particles(600, 3, dt)
    sc, sp = _params(lib, e)
    n = len(pos)
    acc = np.zeros_like(pos)
    pc = PolygonCells(sp, vrts, tris, lib)
    cell = pc.attach(GridBuffers(n, sp.NumCells, lib).cell())
    dpos, dvel, dacc, ddens = DevBuf.from_array(pos), DevBuf.from_array(vel), DevBuf.from_array(acc), DevBuf(4 * n)
    lib.CuSphIntegrateWithPolygon(dpos.ptr, dvel.ptr, dacc.ptr, ddens.ptr, None, pc.d_vrts.ptr, pc.d_tris.ptr, len(tris), dt, cell)
    check(lib)
    gp, gv = dpos.download(np.float32).reshape(n, 4), dvel.download(np.float32).reshape(n, 4)
    P = oracle.params_from(sp, sc)
    rp, rv = pos.copy(), vel.copy()
    oracle.load().orc_integrate_polygon_f(oracle.ptr(pos), oracle.ptr(vel), oracle.ptr(acc), None, n, C.byref(P),
                                          oracle.ptr(pc.vrts), oracle.ptr(pc.tris), len(tris), oracle.ptr(pc.start),
                                          oracle.ptr(pc.end), oracle.ptr(pc.index), float(e), dt, oracle.ptr(rp), oracle.ptr(rv))
    assert np.abs(gp[:, :3] - rp[:, :3]).max() <= 1e-5
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-4 * np.abs(rv[:, :3]).max()
    # w components are kept (in place), the crossing particles are stopped in front of the square
    assert np.all(gp[:, 3] == 1.0) and np.all(gv[:, 3] == 2.0)
    _, _, o, e1, e2, nrm, L = _tilted_square()
    h = (gp[:, :3] - o) @ nrm
    assert np.all(h[kind == 0] >= -1e-6)
    assert np.all(h[kind == 1] >= -1e-6) and np.all(gv[kind == 1, :3] @ nrm > 0)  # from behind: passes through


def test_pbf_integrate_with_polygon_into_new_arrays(lib):
    dt = np.float32(1.0 / 120.0)
    pos, vel, vrts, tris, kind = _particles(400, 5, dt)
    sc, sp = _params(lib, 0.2)
    n = len(pos)
    acc = np.zeros_like(pos)
    acc[:, 1] = -9.8
    pc = PolygonCells(sp, vrts, tris, lib)
    cell = pc.attach(GridBuffers(n, sp.NumCells, lib).cell())
    dpos, dvel, dacc = DevBuf.from_array(pos), DevBuf.from_array(vel), DevBuf.from_array(acc)
    dnp, dnv = DevBuf(16 * n), DevBuf(16 * n)
    lib.CuPbdSphIntegrateWithPolygon(dpos.ptr, dvel.ptr, dacc.ptr, None, dnp.ptr, dnv.ptr, pc.d_vrts.ptr, pc.d_tris.ptr, len(tris),
                                     dt, cell)
    check(lib)
    gp, gv = dnp.download(np.float32).reshape(n, 4), dnv.download(np.float32).reshape(n, 4)
    P = oracle.params_from(sp, sc)
    rp, rv = np.zeros_like(pos), np.zeros_like(vel)
    oracle.load().orc_integrate_polygon_f(oracle.ptr(pos), oracle.ptr(vel), oracle.ptr(acc), None, n, C.byref(P),
                                          oracle.ptr(pc.vrts), oracle.ptr(pc.tris), len(tris), oracle.ptr(pc.start),
                                          oracle.ptr(pc.end), oracle.ptr(pc.index), 0.2, dt, oracle.ptr(rp), oracle.ptr(rv))
    assert np.abs(gp[:, :3] - rp[:, :3]).max() <= 1e-5
    assert np.abs(gv[:, :3] - rv[:, :3]).max() <= 1e-4 * np.abs(rv[:, :3]).max()
    assert np.array_equal(dpos.download(np.float32).reshape(n, 4), pos)  # inputs untouched


def test_lattice_densities(lib):
    sc = scene.pbf_dam_break(10, 10, 10)
    pos, vel = sc.particles()
    sp = sc.params()
    lib.CuSetParameters(C.byref(sp))
    n = sc.n
    g = GridBuffers(n, sp.NumCells, lib)
    dpos, dvel = DevBuf.from_array(pos), DevBuf.from_array(vel)
    c = g.set_object_to_cell(dpos.ptr, dvel.ptr)
    nx, ny, nz = 23, 17, 13
    x0, y0, z0, d = np.float32(-0.05), np.float32(-0.04), np.float32(-0.03), np.float32(0.047)
    P = oracle.params_from(sp, sc)
    orc = oracle.load()
    ref = np.zeros(nx * ny * nz, np.float32)
    orc.orc_grid_density_f(oracle.ptr(pos), n, C.byref(P), nx, ny, nz, x0, y0, z0, d, d, d, None, oracle.ptr(ref))
    out = DevBuf(4 * nx * ny * nz)
    for fn in (lib.CuSphGridDensity, lib.CuPbdSphGridDensity):
        out.fill(0)
        fn(out.ptr, c, nx, ny, nz, x0, y0, z0, d, d, d)
        check(lib)
        got = out.download(np.float32)
        assert np.abs(got - ref).max() <= 1e-5 * ref.max()
        assert (ref > 0).sum() > 0.3 * ref.size and (ref == 0).sum() > 0
    flag = np.where(np.arange(n) % 4 == 0, -1.0, 1.0).astype(np.float32)
    dflag = DevBuf.from_array(flag)
    lib.CuIceMeshMake(out.ptr, c, nx, ny, nz, x0, y0, z0, d, d, d, dflag.ptr)
    check(lib)
    reff = np.zeros_like(ref)
    orc.orc_grid_density_f(oracle.ptr(pos), n, C.byref(P), nx, ny, nz, x0, y0, z0, d, d, d, oracle.ptr(flag), oracle.ptr(reff))
    got = out.download(np.float32)
    assert np.abs(got - reff).max() <= 1e-5 * ref.max()
    assert reff.sum() < ref.sum()
