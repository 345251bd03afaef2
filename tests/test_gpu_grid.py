"""Neighbour grid through the C ABI vs the oracle: bit-exact hashes, stable permutation and cell bounds."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import check, scene
from fltk_sph_ice_amd.abi_driver import GridBuffers
from fltk_sph_ice_amd.device import DevBuf

pytestmark = pytest.mark.gpu


def _setup(lib, sc):
    sp = sc.params()
    lib.CuSetParameters(C.byref(sp))
    return sp


def _oracle_grid(pos, sp, attr=None):
    P = oracle.params_from(sp)
    orc = oracle.load()
    n = pos.shape[0]
    h = np.zeros(n, np.uint32)
    idx = np.zeros(n, np.uint32)
    orc.orc_calc_hash_f(oracle.ptr(pos), None if attr is None else oracle.ptr(attr), n, C.byref(P), oracle.ptr(h), oracle.ptr(idx))
    hs, ids = h.copy(), idx.copy()
    orc.orc_stable_sort(oracle.ptr(hs), oracle.ptr(ids), n)
    cs = np.zeros(sp.NumCells, np.uint32)
    ce = np.zeros(sp.NumCells, np.uint32)
    orc.orc_cell_bounds(oracle.ptr(hs), n, sp.NumCells, oracle.ptr(cs), oracle.ptr(ce))
    return h, hs, ids, cs, ce


@pytest.mark.parametrize("dims,seed", [((32, 32, 32), 1), ((17, 9, 5), 7), ((1, 1, 1), 3)])
def test_hash_sort_reorder_bit_exact(lib, dims, seed):
    sc = scene.wcsph_dam_break(*dims, seed=seed)
    pos, vel = sc.particles()
    sp = _setup(lib, sc)
    n = sc.n
    g = GridBuffers(n, sp.NumCells, lib)
    dpos, dvel = DevBuf.from_array(pos), DevBuf.from_array(vel)
    lib.CuCalcHash(g.hash.ptr, g.index.ptr, dpos.ptr, None, n)
    check(lib)
    h_raw = g.hash.download(np.uint32)
    ref_h, ref_hs, ref_ids, ref_cs, ref_ce = _oracle_grid(pos, sp)
    assert np.array_equal(h_raw, ref_h)
    assert np.array_equal(g.index.download(np.uint32), np.arange(n, dtype=np.uint32))
    lib.CuSort(g.hash.ptr, g.index.ptr, n)
    check(lib)
    assert np.array_equal(g.hash.download(np.uint32), ref_hs)
    assert np.array_equal(g.index.download(np.uint32), ref_ids)
    lib.CuReorderDataAndFindCellStart(g.cell(), dpos.ptr, dvel.ptr)
    check(lib)
    assert np.array_equal(g.cstart.download(np.uint32), ref_cs)
    assert np.array_equal(g.cend.download(np.uint32), ref_ce)
    assert np.array_equal(g.spos.download(np.float32).reshape(n, 4), pos[ref_ids])
    assert np.array_equal(g.svel.download(np.float32).reshape(n, 4), vel[ref_ids])


def test_points_on_cell_faces_and_outside_grid(lib):
    """Coordinates exactly on faces, one ulp either side, and outside the world (clamped)."""
    sc = scene.wcsph_dam_break(4, 4, 4)
    sp = _setup(lib, sc)
    o = np.float32(sp.WorldOrigin[0])
    cw = np.float32(sp.CellWidth[0])
    faces = o + cw * np.arange(0, 60, dtype=np.float32)
    vals = np.concatenate([faces, np.nextafter(faces, np.float32(-10)), np.nextafter(faces, np.float32(10)),
                           np.float32([-5.0, 7.0, np.float32(sp.WorldMax[0])])])
    rng = np.random.default_rng(11)
    n = 3000
    pos = np.zeros((n, 4), np.float32)
    for k in range(3):
        pos[:, k] = rng.choice(vals, n)
    g = GridBuffers(n, sp.NumCells, lib)
    dpos = DevBuf.from_array(pos)
    lib.CuCalcHash(g.hash.ptr, g.index.ptr, dpos.ptr, None, n)
    check(lib)
    ref_h = _oracle_grid(pos, sp)[0]
    assert np.array_equal(g.hash.download(np.uint32), ref_h)
    assert ref_h.max() < sp.NumCells


def test_deleted_particles_sort_last_stably(lib):
    sc = scene.wcsph_dam_break(16, 16, 16, seed=4)
    pos, vel = sc.particles()
    sp = _setup(lib, sc)
    n = sc.n
    attr = np.zeros(n, np.int32)
    attr[np.random.default_rng(5).choice(n, n // 3, replace=False)] = -1
    g = GridBuffers(n, sp.NumCells, lib)
    dpos, dattr = DevBuf.from_array(pos), DevBuf.from_array(attr)
    lib.CuCalcHash(g.hash.ptr, g.index.ptr, dpos.ptr, dattr.ptr, n)
    lib.CuSort(g.hash.ptr, g.index.ptr, n)
    check(lib)
    ref_h, ref_hs, ref_ids, _, _ = _oracle_grid(pos, sp, attr)
    assert (ref_h == 2147483647).sum() == n // 3
    assert np.array_equal(g.hash.download(np.uint32), ref_hs)
    assert np.array_equal(g.index.download(np.uint32), ref_ids)


def test_sort_many_equal_keys_is_stable(lib):
    sc = scene.wcsph_dam_break(2, 2, 2)
    sp = _setup(lib, sc)
    n = 70000
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 5, n).astype(np.uint32)  # 5 huge buckets
    vals = rng.permutation(n).astype(np.uint32)
    dk, dv = DevBuf.from_array(keys), DevBuf.from_array(vals)
    lib.CuSort(dk.ptr, dv.ptr, n)
    check(lib)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(dk.download(np.uint32), keys[order])
    assert np.array_equal(dv.download(np.uint32), vals[order])


def test_sort_rejects_foreign_keys(lib):
    sc = scene.wcsph_dam_break(2, 2, 2)
    sp = _setup(lib, sc)
    keys = np.array([1, sp.NumCells + 5, 0], np.uint32)
    dk, dv = DevBuf.from_array(keys), DevBuf.from_array(np.arange(3, dtype=np.uint32))
    lib.CuSort(dk.ptr, dv.ptr, 3)
    with pytest.raises(RuntimeError, match="CuSort"):
        check(lib)


@pytest.mark.parametrize("n", [1, 100, 2048, 2049, 1 << 20, 3_000_001])
def test_scan_uint_and_float(lib, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 100, n).astype(np.uint32)
    da, do = DevBuf.from_array(a), DevBuf(4 * n)
    lib.CuScan(do.ptr, da.ptr, n)
    check(lib)
    ref = np.concatenate([[0], np.cumsum(a.astype(np.uint64))[:-1]]).astype(np.uint32)
    assert np.array_equal(do.download(np.uint32), ref)
    f = rng.random(n).astype(np.float32)
    df = DevBuf.from_array(f)
    lib.CuScanf(do.ptr, df.ptr, n)
    check(lib)
    reff = np.concatenate([[0.0], np.cumsum(f.astype(np.float64))[:-1]])
    assert np.allclose(do.download(np.float32), reff, rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("n", [1 << 20, 4097, 4096, 1])
def test_sort_all_keys_in_one_bucket(lib, n):
    # the stable sort is linear whatever the key distribution: every key equal keeps the input order exactly
    sc = scene.wcsph_dam_break(2, 2, 2)
    sp = _setup(lib, sc)
    keys = np.full(n, 3, np.uint32)
    vals = np.random.default_rng(1).permutation(n).astype(np.uint32)
    dk, dv = DevBuf.from_array(keys), DevBuf.from_array(vals)
    lib.CuSort(dk.ptr, dv.ptr, n)
    check(lib)
    assert np.array_equal(dk.download(np.uint32), keys)
    assert np.array_equal(dv.download(np.uint32), vals)


def test_sort_border_cell_with_a_tenth_of_the_particles(lib):
    # 10% of 1M particles clamped into the corner cell of the grid (particles outside the world), the rest spread
    sc = scene.wcsph_dam_break(8, 8, 8)
    sp = _setup(lib, sc)
    n = 1 << 20
    rng = np.random.default_rng(5)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    out = rng.choice(n, n // 10, replace=False)
    pos[out, :3] = rng.uniform(-3.0, -1.0, (len(out), 3)).astype(np.float32)  # all hash to cell 0
    g = GridBuffers(n, sp.NumCells, lib)
    dpos = DevBuf.from_array(pos)
    lib.CuCalcHash(g.hash.ptr, g.index.ptr, dpos.ptr, None, n)
    lib.CuSort(g.hash.ptr, g.index.ptr, n)
    check(lib)
    ref_h, ref_hs, ref_ids, _, _ = _oracle_grid(pos, sp)
    assert (ref_h == 0).sum() >= n // 10
    assert np.array_equal(g.hash.download(np.uint32), ref_hs)
    assert np.array_equal(g.index.download(np.uint32), ref_ids)


@pytest.mark.parametrize("nbits_keys", [(1000, 200), (5000, 300), (200000, 1 << 21), (70001, (1 << 24) - 1)])
def test_sort_random_keys_matches_stable_argsort(lib, nbits_keys):
    # one-, two- and three-pass key widths (bucket count = NumCells of a grid with that many cells)
    n, ncells = nbits_keys
    sc = scene.wcsph_dam_break(2, 2, 2)
    sp = sc.params()
    sp.NumCells = ncells
    lib.CuSetParameters(C.byref(sp))
    rng = np.random.default_rng(n)
    keys = rng.integers(0, ncells, n).astype(np.uint32)
    keys[::97] = 2147483647  # deleted-particle markers sort last
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = DevBuf.from_array(keys), DevBuf.from_array(vals)
    lib.CuSort(dk.ptr, dv.ptr, n)
    check(lib)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(dk.download(np.uint32), keys[order])
    assert np.array_equal(dv.download(np.uint32), vals[order])
