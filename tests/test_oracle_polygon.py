"""Pins the oracle's polygon-obstacle integrator (rx_sph_cpu.cpp:956-1077 with the segment/triangle test of
rx_cu_common.cu:891-960) and its lattice density (rx_sph_kernel.cu:1039-1126) with hand-worked collisions, an
independent point-in-triangle test and the particle density of the SPH path."""
import ctypes as C

import numpy as np
import pytest

from oracle import ptr
from test_oracle_pins import lattice, make_params, pos4


def _cells_all(P, ntri):
    """Every triangle listed in every cell (the collision test alone decides)."""
    nc = P.gsize[0] * P.gsize[1] * P.gsize[2]
    start = (np.arange(nc, dtype=np.uint32) * ntri).astype(np.uint32)
    end = (start + ntri).astype(np.uint32)
    idx = np.tile(np.arange(ntri, dtype=np.uint32), nc)
    return start, end, idx


def _square(y=0.5, lo=0.2, hi=0.8, up=True):
    """Two triangles covering [lo, hi]^2 in x-z at height y, normal +y (or -y)."""
    v = np.array([[lo, y, lo], [hi, y, lo], [hi, y, hi], [lo, y, hi]], np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2]] if up else [[0, 1, 2], [0, 2, 3]], np.int32)
    return v, t


def _run(orc, P, p, v, vrts, tris, e, dt, cells=None, attr=None):
    n = len(p)
    acc = np.zeros_like(p)
    start, end, idx = cells if cells is not None else _cells_all(P, len(tris))
    op, ov = np.zeros_like(p), np.zeros_like(v)
    orc.orc_integrate_polygon_f(ptr(p), ptr(v), ptr(acc), None if attr is None else ptr(attr), n, C.byref(P), ptr(vrts),
                                ptr(tris), len(tris), ptr(start), ptr(end), ptr(idx), e, np.float32(dt), ptr(op), ptr(ov))
    return op, ov


def test_normal_incidence_mirrors_the_penetration(orc):
    # falling through y = 0.5 at speed 3 with dt = 0.01: d = 0.01, l = 0.03.  e = 0: v_y -> 0 and the point goes
    # to the crossing point plus (dt d / l) * 3 = d above the plane.  e = 0.6: e' = e d/(dt |v|) = 0.2,
    # v_y -> 3 e' = 0.6 and the point goes 1.2 d above the plane.
    P = make_params(lo=0.0, hi=1.0)
    vrts, tris = _square()
    p = pos4([[0.4, 0.52, 0.6]])
    v = pos4([[0.0, -3.0, 0.0]])
    op, ov = _run(orc, P, p, v, vrts, tris, 0.0, 0.01)
    assert op[0, 1] == pytest.approx(0.51, abs=1e-6) and ov[0, 1] == pytest.approx(0.0, abs=1e-5)
    assert op[0, 0] == pytest.approx(0.4) and op[0, 2] == pytest.approx(0.6)
    op, ov = _run(orc, P, p, v, vrts, tris, 0.6, 0.01)
    assert ov[0, 1] == pytest.approx(0.6, rel=1e-5)
    assert op[0, 1] == pytest.approx(0.5 + 1.2 * 0.01, abs=1e-6)


def test_oblique_hit_keeps_the_tangential_velocity(orc):
    P = make_params(lo=0.0, hi=1.0)
    vrts, tris = _square()
    p = pos4([[0.4, 0.52, 0.6]])
    v = pos4([[2.0, -3.0, -1.0]])
    op, ov = _run(orc, P, p, v, vrts, tris, 0.0, 0.01)
    assert ov[0, 0] == pytest.approx(2.0) and ov[0, 2] == pytest.approx(-1.0) and ov[0, 1] == pytest.approx(0.0, abs=1e-5)
    assert op[0, 1] > 0.5


def test_back_side_and_misses_pass_through(orc):
    P = make_params(lo=0.0, hi=1.0)
    vrts, tris = _square()
    # from below (against the normal), outside the square, and not reaching the plane
    p = pos4([[0.4, 0.49, 0.6], [0.9, 0.52, 0.9], [0.4, 0.56, 0.6]])
    v = pos4([[0.0, 3.0, 0.0], [0.0, -3.0, 0.0], [0.0, -3.0, 0.0]])
    op, ov = _run(orc, P, p, v, vrts, tris, 0.5, 0.01)
    np.testing.assert_allclose(op[:, :3], p[:, :3] + 0.01 * v[:, :3], rtol=1e-6)
    np.testing.assert_array_equal(ov, v)
    # a downward-facing square is hit from below: 0.49 -> 0.52 penetrates by 0.02, mirrored to 0.48
    vd, td = _square(up=False)
    op, ov = _run(orc, P, p[:1], v[:1], vd, td, 0.0, 0.01)
    assert op[0, 1] == pytest.approx(0.48, abs=1e-6) and ov[0, 1] == pytest.approx(0.0, abs=1e-5)


def test_hit_region_is_the_triangle(orc):
    # one triangle in the plane y = 0.5; vertical segments hit it exactly where the crossing point lies inside
    # (sign-of-cross-product test, not the parametric one of the oracle)
    P = make_params(lo=0.0, hi=1.0)
    vrts = np.array([[0.2, 0.5, 0.3], [0.7, 0.5, 0.8], [0.8, 0.5, 0.2]], np.float32)
    a, b, c = vrts[0, [0, 2]], vrts[1, [0, 2]], vrts[2, [0, 2]]
    n_up = np.cross(vrts[1] - vrts[0], vrts[2] - vrts[0])
    tris = np.array([[0, 1, 2]] if n_up[1] > 0 else [[0, 2, 1]], np.int32)
    rng = np.random.default_rng(3)
    xz = rng.uniform(0.15, 0.85, (400, 2))

    def side(p, q, r):
        return (q[0] - p[0]) * (r[:, 1] - p[1]) - (q[1] - p[1]) * (r[:, 0] - p[0])

    s1, s2, s3 = side(a, b, xz), side(b, c, xz), side(c, a, xz)
    inside = ((s1 > 0) & (s2 > 0) & (s3 > 0)) | ((s1 < 0) & (s2 < 0) & (s3 < 0))
    margin = np.minimum(np.minimum(np.abs(s1), np.abs(s2)), np.abs(s3)) > 1e-4
    p = pos4(np.column_stack([xz[:, 0], np.full(len(xz), 0.52), xz[:, 1]]))
    v = pos4(np.tile([0.0, -3.0, 0.0], (len(xz), 1)))
    op, ov = _run(orc, P, p, v, vrts, tris, 0.0, 0.01)
    hit = op[:, 1] > 0.5
    assert inside.sum() > 50 and (~inside).sum() > 50
    assert np.array_equal(hit[margin], inside[margin])


def test_only_listed_cells_are_tested_and_attr_skips(orc):
    P = make_params(lo=0.0, hi=1.0)
    vrts, tris = _square()
    p = pos4([[0.4, 0.52, 0.6], [0.4, 0.52, 0.6]])
    v = pos4([[0.0, -3.0, 0.0], [0.0, -3.0, 0.0]])
    nc = P.gsize[0] * P.gsize[1] * P.gsize[2]
    empty = (np.full(nc, 0xFFFFFFFF, np.uint32), np.full(nc, 0xFFFFFFFF, np.uint32), np.zeros(1, np.uint32))
    op, _ = _run(orc, P, p, v, vrts, tris, 0.0, 0.01, cells=empty)
    assert np.allclose(op[:, 1], 0.49)
    attr = np.array([0, -1], np.int32)
    op, ov = _run(orc, P, p, v, vrts, tris, 0.0, 0.01, attr=attr)
    assert op[0, 1] == pytest.approx(0.51, abs=1e-6)
    np.testing.assert_array_equal(op[1], p[1])
    np.testing.assert_array_equal(ov[1], v[1])


def test_lattice_density_at_particles_equals_particle_density(orc):
    # lattice nodes placed on the particles of a lattice block: node density = SPH density of that particle
    s, h = 0.05, 0.1
    P = make_params(h=h, lo=0.0, hi=1.0)
    x = lattice(8, s, (0.3, 0.3, 0.3))
    p = pos4(x)
    n = len(p)
    dens, pres = np.zeros(n, np.float32), np.zeros(n, np.float32)
    orc.orc_sph_density_f(ptr(p), n, C.byref(P), ptr(dens), ptr(pres))
    g = np.zeros(8 ** 3, np.float32)
    orc.orc_grid_density_f(ptr(p), n, C.byref(P), 8, 8, 8, np.float32(0.3), np.float32(0.3), np.float32(0.3), np.float32(s),
                           np.float32(s), np.float32(s), None, ptr(g))
    # lattice() orders x slowest; the node array orders x fastest
    gi = g.reshape(8, 8, 8).transpose(2, 1, 0).ravel()
    np.testing.assert_allclose(gi, dens, rtol=1e-5)  # nodes and particles may differ by an ulp of position
    # the ice flag leaves out the flagged particles: same as the density of the others alone
    flag = np.where(np.arange(n) % 3 == 0, -1.0, 1.0).astype(np.float32)
    gf = np.zeros_like(g)
    orc.orc_grid_density_f(ptr(p), n, C.byref(P), 8, 8, 8, np.float32(0.3), np.float32(0.3), np.float32(0.3), np.float32(s),
                           np.float32(s), np.float32(s), ptr(flag), ptr(gf))
    keep = np.ascontiguousarray(p[flag >= 0])
    gk = np.zeros_like(g)
    orc.orc_grid_density_f(ptr(keep), len(keep), C.byref(P), 8, 8, 8, np.float32(0.3), np.float32(0.3), np.float32(0.3),
                           np.float32(s), np.float32(s), np.float32(s), None, ptr(gk))
    np.testing.assert_array_equal(gf, gk)
    assert gf.sum() < g.sum()
