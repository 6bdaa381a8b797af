"""The oracle's restatement of glibc rand() (the generator of the reference harness's inputs) against
the C library's own rand(), and the reference's core order and entry expression."""
import ctypes
import ctypes.util

import numpy as np
import pytest

from oracle import glibc_rand


def _libc():
    name = ctypes.util.find_library("c")
    if not name:
        pytest.skip("no C library found")
    libc = ctypes.CDLL(name)
    libc.srand.argtypes = [ctypes.c_uint]
    libc.rand.restype = ctypes.c_int
    return libc


@pytest.mark.parametrize("seed", [1134, 1, 0, 12345, 2 ** 31 + 5])
def test_rand_stream_matches_libc(seed):
    libc = _libc()
    libc.srand(seed)
    want = np.array([libc.rand() for _ in range(5000)], dtype=np.uint32)
    assert np.array_equal(glibc_rand.rand_stream(seed, 5000), want)


def test_uniform_expression():
    # (float)(v / RAND_MAX) * 2.0 - 1.0: the ends of the range and the float rounding of the quotient
    v = np.array([0, glibc_rand.RAND_MAX, glibc_rand.RAND_MAX // 2, 1, 123456789], dtype=np.uint32)
    u = glibc_rand.uniform(v)
    assert u[0] == -1.0 and u[1] == 1.0
    for x, y in zip(v, u):
        q = np.float32(float(x) / 2147483647.0)
        assert y == np.float32(float(q) * 2.0 - 1.0)
    assert np.all(np.abs(glibc_rand.uniform(glibc_rand.rand_stream(1134, 10000))) <= 1.0)


def test_ref_cores_order_and_sizes():
    """Order 6: G5 and G6 take the first values, then G4, G3, G2, G1 (ttdec6_singleGPU.cu:198-263);
    TR sizes: G1 N1 r0 r1, G6 r5 N6 r0."""
    dims, ranks = [5, 6, 7, 4, 3, 2], [1, 2, 3, 2, 3, 2, 1]
    cores = glibc_rand.ref_cores(dims, ranks, seed=1134)
    sizes = [ranks[k] * dims[k] * ranks[k + 1] for k in range(6)]
    assert [c.size for c in cores] == sizes
    stream = glibc_rand.uniform(glibc_rand.rand_stream(1134, sum(sizes)))
    order = [4, 5, 3, 2, 1, 0]
    pos = 0
    for k in order:
        assert np.array_equal(cores[k], stream[pos:pos + sizes[k]])
        pos += sizes[k]
    tr = glibc_rand.ref_cores(dims, [2, 2, 3, 2, 3, 2, 2], tr=True)
    assert tr[0].size == 5 * 2 * 2 and tr[5].size == 2 * 2 * 2
