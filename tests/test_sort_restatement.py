"""The library's restatement of libstdc++ std::sort (used on the device for the
reference's unstable sorts) yields the permutation of the real std::sort.  The real
sort runs in the oracle library: its findSeeds sorts SMEMs by loglik with std::sort;
here it is driven through a tiny g++-built helper."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import msgseqtk_amd as M

HELPER = r'''
#include <algorithm>
#include <cstdint>
extern "C" void std_sort_perm(const double* k, int32_t* p, int64_t n) {
    std::sort(p, p + n, [&](int32_t a, int32_t b) { return k[a] < k[b]; });
}
'''


@pytest.fixture(scope="module")
def stdsort():
    d = tempfile.mkdtemp()
    src, so = os.path.join(d, "h.cpp"), os.path.join(d, "h.so")
    open(src, "w").write(HELPER)
    subprocess.check_call(["g++", "-O2", "-shared", "-fPIC", src, "-o", so])
    L = ctypes.CDLL(so)
    L.std_sort_perm.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32), ctypes.c_int64]
    return L


def _std(L, keys):
    keys = np.ascontiguousarray(keys, np.float64)
    p = np.arange(len(keys), dtype=np.int32)
    L.std_sort_perm(keys.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                    p.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(keys))
    return p


@pytest.mark.parametrize("n", [0, 1, 2, 5, 16, 17, 33, 100, 1000, 5000])
def test_sort_perm_matches_std_sort(stdsort, n):
    rng = np.random.default_rng(n)
    for trial in range(20):
        keys = rng.integers(0, max(1, n // 4 + 1), size=n).astype(np.float64)   # many ties
        assert np.array_equal(M.sort_by_key(keys), _std(stdsort, keys))


def test_sort_heap_fallback(stdsort):
    """median-of-3 killer sequences drive introsort into its heapsort fallback"""
    for n in (64, 257, 1024, 4096):
        k = n // 2
        keys = np.zeros(n)
        for i in range(1, k + 1):
            if i % 2:
                keys[i - 1] = i
                keys[i] = k + i
            keys[k + i - 1] = 2 * i
        keys = keys // 3    # ties as well
        assert np.array_equal(M.sort_by_key(keys), _std(stdsort, keys))
