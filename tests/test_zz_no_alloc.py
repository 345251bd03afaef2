"""CuSortNoAlloc (the allocation-free CuSort, static scratch) through the C ABI against CuSort: two calls in a row,
the second with a smaller input, give CuSort's keys and values exactly; sizes beyond the static scratch are
rejected before any work."""
import ctypes as C

import numpy as np
import pytest

from fltk_sph_ice_amd import check, load, scene
from fltk_sph_ice_amd.device import DevBuf

pytestmark = pytest.mark.gpu
MAX_N = 1 << 21


def _keys(n, ncells, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, ncells, n).astype(np.uint32)
    k[rng.choice(n, max(1, n // 50), replace=False)] = 2147483647  # deleted-particle markers
    k[: n // 10] = 7  # a crowded bucket
    return k, rng.permutation(n).astype(np.uint32)


def _sort(lib, fn, keys, vals):
    dk, dv = DevBuf.from_array(keys), DevBuf.from_array(vals)
    fn(dk.ptr, dv.ptr, len(keys))
    check(lib)
    return dk.download(np.uint32), dv.download(np.uint32)


def test_no_alloc_sort_matches_cusort_on_two_calls():
    lib = load()
    sc = scene.wcsph_dam_break(32, 32, 32)
    sp = sc.params()
    lib.CuSetParameters(C.byref(sp))
    for n, seed in ((MAX_N, 1), (100_003, 2)):  # the full static size, then a smaller input
        keys, vals = _keys(n, sp.NumCells, seed)
        ref_k, ref_v = _sort(lib, lib.CuSort, keys, vals)
        got_k, got_v = _sort(lib, lib.CuSortNoAlloc, keys, vals)
        assert np.array_equal(got_k, ref_k) and np.array_equal(got_v, ref_v), n
        order = np.argsort(keys, kind="stable")  # and both are the stable order
        assert np.array_equal(got_k, keys[order]) and np.array_equal(got_v, vals[order])


def test_no_alloc_sort_rejects_sizes_beyond_its_scratch():
    lib = load()
    sp = scene.wcsph_dam_break(4, 4, 4).params()
    lib.CuSetParameters(C.byref(sp))
    lib.CuClearLastError()
    lib.CuSortNoAlloc(None, None, MAX_N + 1)  # rejected on the host before any device work
    assert lib.CuGetLastErrorCode() == 1
    assert b"static scratch" in lib.CuGetLastErrorString()
    lib.CuClearLastError()
