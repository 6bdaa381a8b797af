"""The C / OpenMP restatement of the randomized step (oracle/cpu_edrsvd.c), which bench.py times as
the CPU baseline of the third-order configuration, against the FP64 numpy oracle on the same input
and sketch: singular values, cores and reconstruction error."""
import os

import numpy as np
import pytest

from oracle import cpu, ttsvd


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(cpu.LIB):
        from oracle.build_cpu import build
        build()


@pytest.mark.parametrize("dims,ranks,s,q", [
    ([40, 36, 30], [1, 6, 5, 1], 4, 1),
    ([24, 20, 18, 16], [1, 8, 7, 6, 1], 6, 2),
])
def test_cpu_tt_svd_matches_oracle(dims, ranks, s, q):
    A = ttsvd.add_noise(ttsvd.tt_full(ttsvd.make_tt_cores(dims, ranks), dims, ranks).astype(np.float32), 1e-6)
    oms = cpu.omegas(dims, ranks, s)
    cores, sigs = cpu.tt_svd(A, dims, ranks, s, q, oms)
    c_ref, s_ref, p_ref, used = ttsvd.tt_svd(A.astype(np.float64), dims, ranks, s=s, q=q)
    assert all(p == "rsvd" for p in p_ref)
    for k in range(len(dims) - 1):
        r = ranks[k + 1]
        np.testing.assert_allclose(sigs[k][:r], s_ref[k][:r], rtol=0, atol=1e-6 * s_ref[k][0])
        G = cores[k].reshape(-1, r, order="F").astype(np.float64)
        assert np.linalg.norm(G - c_ref[k]) <= 1e-5 * np.linalg.norm(c_ref[k]), k
    shapes = ttsvd.tt_core_shapes(dims, ranks)
    full = ttsvd.tt_full([c.reshape(a, b, order="F") for c, (a, b) in zip(cores, shapes)], dims, ranks)
    e, e_ref = ttsvd.rel_error(A, full), ttsvd.rel_error(A, ttsvd.tt_full(c_ref, dims, used))
    assert abs(e - e_ref) <= 1e-7 + 1e-3 * e_ref, (e, e_ref)


def test_cpu_rsvd_refuses_bad_shapes():
    A = np.zeros(20 * 10, dtype=np.float32)
    with pytest.raises(RuntimeError):
        cpu.rsvd(A, 20, 10, 12, 1, np.zeros((10, 12), np.float32))    # ell > n
