"""The third-order configuration the reference's CPU-path baseline is quoted on: plain TT-SVD in FP32 of
a 256 x 256 x 256 tensor built from seeded cores G1 1x256x16, G2 16x256x16, G3 16x256x1 (N(0, 1)
entries) plus N(0, (1e-6 rms)^2) noise; ranks (16, 16), oversampling 8, 2 power iterations, Gaussian
sketch from the same seed (the EDrSVD of basicRandSVD_LU,
ttdecomposition/third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:557-775, driver :778-1014).
Both steps are randomized (ell = 24 < min(m, n) = 256): A_1 is 256 x 65536, A_2 4096 x 256.

Tolerances (FP32): singular values 1e-5 of sigma_1, cores 1e-4, reconstruction error within
1e-6 + 2e-2 err_oracle of the FP64 oracle's (both at the 1e-6 noise level)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import ttsvd

pytestmark = pytest.mark.gpu

DIMS, RANKS, S, Q = [256, 256, 256], [1, 16, 16, 1], 8, 2
PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "tensor_train_ring_decompositions_hpc_gpu_amd")


@pytest.fixture(scope="module")
def cfg1(engine):
    A = engine.generate(DIMS, RANKS, std=1.0, noise_rel=1e-6, seed=1134)
    An = A.cpu().numpy().astype(np.float64)
    c_ref, s_ref, p_ref, used = ttsvd.tt_svd(An, DIMS, RANKS, s=S, q=Q)
    err_ref = ttsvd.rel_error(An, ttsvd.tt_full(c_ref, DIMS, used))
    return A, An, c_ref, s_ref, p_ref, err_ref


def test_cfg1_tt_svd_fp32_matches_oracle(engine, cfg1):
    A, An, c_ref, s_ref, p_ref, err_ref = cfg1
    assert p_ref == ["rsvd", "rsvd"]
    cores, sig, rep = engine.tt_svd(A, DIMS, RANKS, s=S, q=Q, precision="fp32")
    assert rep["ranks"] == RANKS and rep["path"] == [2, 2] and rep["ell"] == [24, 24]
    assert rep["power_iters"] == [2, 2]
    for k in range(2):
        np.testing.assert_allclose(sig[k][:16], s_ref[k][:16], rtol=0, atol=1e-5 * s_ref[k][0])
        g, shape = cores[k]
        G = g.cpu().numpy()[: shape[0] * shape[1]].reshape(shape, order="F").astype(np.float64)
        assert np.linalg.norm(G - c_ref[k]) <= 1e-4 * np.linalg.norm(c_ref[k]), k
    err = engine.rel_error(cores, DIMS, RANKS, A)
    assert err_ref < 3e-6
    assert abs(err - err_ref) <= 1e-6 + 2e-2 * err_ref, (err, err_ref)


def test_cfg1_through_reference_driver(engine, cfg1, monkeypatch):
    """ttdec_gpu_float (func.h:16) with the engine set to the configuration (FP32, s = 8, q = 2 fixed):
    its singleError is the oracle's error and its cores are the oracle's."""
    A, An, c_ref, s_ref, p_ref, err_ref = cfg1
    path = os.path.join(PKG, "libttx_ttdec6.so")
    if not os.path.exists(path):
        pytest.skip("shim library not built")
    monkeypatch.setenv("TTX_PRECISION", "fp32")
    monkeypatch.setenv("TTX_OVERSAMPLING", str(S))
    monkeypatch.setenv("TTX_POWER_ITERS", str(Q))
    monkeypatch.setenv("TTX_ADAPTIVE", "0")
    lib = ctypes.CDLL(path)
    fn = lib._Z15ttdec_gpu_floatPfPiS0_PdS_S_
    fn.restype = ctypes.c_float
    PF, PI = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
    fn.argtypes = [PF, PI, PI, ctypes.POINTER(ctypes.c_double), PF, PF]
    hA = np.ascontiguousarray(A.cpu().numpy(), dtype=np.float32)
    n3 = (ctypes.c_int * 3)(*DIMS)
    r4 = (ctypes.c_int * 4)(*RANKS)
    eps = (ctypes.c_double * 2)(1e-6, 1e-8)
    err = ctypes.c_float(-1.0)
    ratio = ctypes.c_float(-1.0)
    t = fn(hA.ctypes.data_as(PF), n3, r4, eps, ctypes.byref(err), ctypes.byref(ratio))
    assert t > 0
    assert abs(err.value - err_ref) <= 1e-6 + 2e-2 * err_ref, (err.value, err_ref)
    assert ratio.value == pytest.approx(256 ** 3 / (256 * 16 + 16 * 256 * 16 + 16 * 256), rel=1e-6)
    lib.ttx_last_cores.restype = ctypes.c_int32
    lib.ttx_last_cores.argtypes = [ctypes.POINTER(ctypes.POINTER(ctypes.c_void_p)), ctypes.POINTER(ctypes.c_int32)]
    ptrs = ctypes.POINTER(ctypes.c_void_p)()
    used = (ctypes.c_int32 * 17)()
    assert lib.ttx_last_cores(ctypes.byref(ptrs), used) == 3 and list(used[:4]) == RANKS
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    for k, (a, b) in enumerate([(256, 16), (4096, 16)]):
        h = np.empty(a * b, dtype=np.float32)
        assert hip.hipMemcpy(h.ctypes.data, ptrs[k], 4 * a * b, 2) == 0
        G = h.reshape(a, b, order="F").astype(np.float64)
        assert np.linalg.norm(G - c_ref[k]) <= 1e-4 * np.linalg.norm(c_ref[k]), k
