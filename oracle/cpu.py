"""ctypes driver of the C restatement (cpu_edrsvd.c): the reference's TT sweep of randomized steps
(ttdec_gpu_float, ttdecomposition/third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:778-1014, with
basicRandSVD_LU :557-775 per step) on the host.

TEST INFRASTRUCTURE ONLY: the CPU baseline of bench.py and the tests that check it against the FP64
oracle; the engine never loads it."""
import ctypes
import os

import numpy as np

from . import philox

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcpu_edrsvd.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/libcpu_edrsvd.so is not built (python oracle/build_cpu.py)")
        lib = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        lib.ttx_cpu_rsvd.restype = ctypes.c_int
        lib.ttx_cpu_rsvd.argtypes = [P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int, P, P, P, P]
        lib.ttx_cpu_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def threads():
    return int(load().ttx_cpu_threads())


def rsvd(A, m, n, ell, q, Omega):
    """One randomized step on the column-major m x n float32 A: (U m x ell, sigma, SVT ell x n), both
    float32 column-major (as flat arrays), sigma float64."""
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.float32).reshape(-1)
    Om = np.ascontiguousarray(Omega, dtype=np.float32)
    U = np.empty(m * ell, dtype=np.float32)
    S = np.empty(ell, dtype=np.float64)
    SVT = np.empty(ell * n, dtype=np.float32)
    rc = lib.ttx_cpu_rsvd(A.ctypes.data, m, n, ell, q, Om.ctypes.data, U.ctypes.data, S.ctypes.data, SVT.ctypes.data)
    if rc != 0:
        raise RuntimeError("ttx_cpu_rsvd failed (%d)" % rc)
    return U, S, SVT


def omegas(dims, ranks, s, seed=1234):
    """The sketch of every step (Philox stream k for step k), generated outside the timed region."""
    out = []
    n = int(np.prod(dims))
    for k in range(len(dims) - 1):
        n //= dims[k]
        out.append(philox.omega(seed, k, n, ranks[k + 1] + s))
    return out


def tt_svd(A, dims, ranks, s, q, oms):
    """TT sweep of randomized steps (every step randomized, ell = r + s < min(m, n)); returns cores
    (flat float32, the TT layouts) and the singular values of every step."""
    d = len(dims)
    cur = np.ascontiguousarray(A, dtype=np.float32).reshape(-1)
    n = cur.size
    cores, sigs = [], []
    for k in range(d - 1):
        m = ranks[k] * dims[k]
        n //= dims[k]
        r, ell = ranks[k + 1], ranks[k + 1] + s
        U, S, SVT = rsvd(cur, m, n, ell, q, oms[k])
        cores.append(U[: m * r].copy())
        sigs.append(S)
        cur = np.ascontiguousarray(SVT.reshape(n, ell)[:, :r]).reshape(-1)   # rows 0..r-1 of SVT
    cores.append(cur)
    return cores, sigs
