"""GPU parity of the search and intersection operators against the oracle and numpy."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from gary_amd import ops

pytestmark = pytest.mark.gpu


def _sorted_unique(rng, n, lo, hi):
    return np.unique(rng.integers(lo, hi, size=n)).astype(np.int32)


@pytest.fixture(scope="module")
def search_data():
    # the shape of test/SearchTest.cu:196-210: 4 Mi sorted ints in [2^29, 2^30], 32 Ki keys
    rng = np.random.default_rng(0)
    lst = np.sort(rng.integers(2**29, 2**30, size=4 * 1024 * 1024)).astype(np.int32)
    keys = rng.integers(2**29, 2**30, size=32 * 1024).astype(np.int32)
    keys[::3] = rng.choice(lst, keys[::3].size)
    return lst, keys


@pytest.mark.parametrize("mode", [ops.SEARCH_PLAIN, ops.SEARCH_2PHASE])
@pytest.mark.parametrize("sort_keys", [True, False])
def test_search_reference_shape(search_data, mode, sort_keys):
    lst, keys = search_data
    if sort_keys:
        keys = np.sort(keys)
    want = np.isin(keys, lst)  # == std::binary_search (the reference's oracle)
    for cd in (32, 64, 256):
        got = ops.binary_search(lst, keys, mode, cd)
        assert np.array_equal(got, want), cd
        assert ops.binary_search_count(lst, keys, mode, cd) == int(want.sum())


@pytest.mark.parametrize("cd", [1, 8, 32, 64, 128, 1024])
def test_search_2phase_small_lists(cd):
    rng = np.random.default_rng(cd)
    for size in (1, 2, 3, 31, 32, 33, 64, 65, 200, 1000):
        lst = np.sort(rng.choice(np.arange(0, 4 * size + 8), size=size, replace=False)).astype(np.int32)
        keys = np.arange(int(lst[0]) - 1, int(lst[-1]) + 2, dtype=np.int32)
        cache = oracle.build_cache(lst, cd)
        want = np.array([oracle.binary_search_2phase(lst, cache, k) for k in keys])
        assert np.array_equal(want, np.isin(keys, lst))
        assert np.array_equal(ops.binary_search(lst, keys, ops.SEARCH_2PHASE, cd), want), size


def test_search_empty():
    assert ops.binary_search(np.zeros(0, np.int32), np.arange(5)).sum() == 0
    assert ops.binary_search_count(np.arange(10), np.zeros(0, np.int32)) == 0


@pytest.mark.parametrize("mode", [ops.SEARCH_PLAIN, ops.SEARCH_2PHASE])
def test_batched_search(mode):
    # batchedBinarySearch{,2Phase}Kernel (test/SearchTest.cu:543-644), scaled down
    rng = np.random.default_rng(7)
    lists, keys = [], []
    for i in range(37):
        n = int(rng.integers(0 if i == 5 else 1, 20000))
        lists.append(np.sort(rng.integers(0, 1 << 20, size=n)).astype(np.int32))
        k = rng.integers(0, 1 << 20, size=int(rng.integers(0, 3000))).astype(np.int32)
        if n and k.size:
            k[::2] = rng.choice(lists[-1], k[::2].size)
        keys.append(k)
    got = ops.batched_binary_search(lists, keys, mode, 64)
    for l, k, g in zip(lists, keys, got):
        assert np.array_equal(g, np.isin(k, l))


@pytest.mark.parametrize("method", [ops.ISECT_BSEARCH, ops.ISECT_CACHED_BSEARCH, ops.ISECT_MERGE,
                                    ops.ISECT_BITMAP])
def test_intersect_count(method):
    rng = np.random.default_rng(11)
    # the shape of test/SetOpTest.cu:62-68 (4 Mi vs 32 Ki in [2^26, 2^30])
    a = _sorted_unique(rng, 4 * 1024 * 1024, 2**26, 2**30)
    b = _sorted_unique(rng, 32 * 1024, 2**26, 2**30)
    b[::4] = rng.choice(a, b[::4].size)
    b = np.unique(b)
    want = oracle.intersect_count(a, b)
    assert want == np.intersect1d(a, b).size
    assert ops.intersect_count(a, b, method, 64) == want
    assert ops.intersect_count(b, a, method, 64) == want
    for trial in range(40):
        na, nb = rng.integers(0, 3000, size=2)
        x = _sorted_unique(rng, na, 0, 5000)
        y = _sorted_unique(rng, nb, 0, 5000)
        assert ops.intersect_count(x, y, method, 32) == oracle.intersect_count(x, y), trial


@pytest.mark.parametrize("method", [ops.ISECT_BSEARCH, ops.ISECT_CACHED_BSEARCH])
def test_intersect_emit_sorted(method):
    rng = np.random.default_rng(3)
    a = _sorted_unique(rng, 1 << 20, 0, 1 << 24)
    b = _sorted_unique(rng, 1 << 15, 0, 1 << 24)
    b[::2] = rng.choice(a, b[::2].size)
    b = np.unique(b)
    want = oracle.intersect_emit(a, b)
    assert np.array_equal(want, np.intersect1d(a, b))
    assert np.array_equal(ops.intersect(a, b, method, 64), want)
    assert np.array_equal(ops.intersect(b, a, method, 64), want)
    assert ops.intersect(np.zeros(0, np.int32), b, method).size == 0
