"""Reference-name shims (include/ttx_shim.h): the exact C++ symbols of ttdec_half_lu_device_6 /
ttdec_half_lu_host_6 (ttdec6_singleGPU.cu:656,799; trdec6_singleGPU.cu:668,827), callable the way the
reference's main() calls them, with the cores reachable through ttx_last_cores."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import ttsvd

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "tensor_train_ring_decompositions_hpc_gpu_amd")
DEVICE6 = "_Z22ttdec_half_lu_device_6PfPKiS1_RS0_S2_b"
HOST6 = "_Z20ttdec_half_lu_host_6PfPiPKiRS1_S3_b"
LIBS = {"tt": "libttx_ttdec6.so", "tr": "libttx_trdec6.so"}


def _load(kind):
    path = os.path.join(PKG, LIBS[kind])
    if not os.path.exists(path):
        pytest.skip("shim library not built")
    lib = ctypes.CDLL(path)
    PI = ctypes.POINTER(ctypes.c_int)
    for name in (DEVICE6, HOST6):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_float
        fn.argtypes = [ctypes.POINTER(ctypes.c_float), PI, PI, PI, PI, ctypes.c_bool]
    lib.ttx_last_cores.restype = ctypes.c_int32
    lib.ttx_last_cores.argtypes = [ctypes.POINTER(ctypes.POINTER(ctypes.c_void_p)), ctypes.POINTER(ctypes.c_int32)]
    lib.ttx_last_rel_error.restype = ctypes.c_double
    return lib


@pytest.mark.parametrize("kind", ["tt", "tr"])
def test_shim_symbols_demangle_to_reference_signatures(kind):
    lib = _load(kind)
    assert getattr(lib, DEVICE6) and getattr(lib, HOST6)
    cxxfilt = shutil.which("c++filt")
    if cxxfilt is None:
        pytest.skip("c++filt not available")
    out = subprocess.run([cxxfilt, DEVICE6, HOST6], stdout=subprocess.PIPE, text=True).stdout.split("\n")
    assert out[0] == "ttdec_half_lu_device_6(float*, int const*, int const*, int const&, int const&, bool)"
    assert out[1] == "ttdec_half_lu_host_6(float*, int*, int const*, int const&, int const&, bool)"


def _call(lib, name, A, ranks, dims, times=1, cal_error=True):
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    r = (ctypes.c_int * 7)(*ranks)
    d = (ctypes.c_int * 6)(*dims)
    hA = np.ascontiguousarray(A, dtype=np.float32)
    t = getattr(lib, name)(hA.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), r, d, ci(times), ci(64), cal_error)
    return t, hA


@pytest.mark.gpu
@pytest.mark.parametrize("kind,name", [("tt", DEVICE6), ("tt", HOST6), ("tr", DEVICE6), ("tr", HOST6)])
def test_shim_runs_like_reference_main(engine, kind, name):
    lib = _load(kind)
    dims = [8, 8, 8, 8, 8, 8]
    if kind == "tt":
        ranks = [1, 4, 4, 4, 4, 4, 1]
        A = engine.generate(dims, ranks, std=1.0, noise_rel=1e-6, seed=1134)
    else:
        # the sequential TR-SVD reproduces a ring of ranks r_k exactly once its inner ranks reach
        # r0^2 r_k: the first-step basis mixes the ring index into the chain index
        ranks = [2, 2, 8, 8, 8, 8, 2]
        A = engine.generate(dims, [2] * 7, std=1.0, noise_rel=1e-6, seed=1134, tr=True)
    hA = A.cpu().numpy()
    before = hA.copy()
    t, _ = _call(lib, name, hA, ranks, dims, times=2)
    assert t > 0
    assert np.array_equal(hA, before)                      # h_A is not overwritten
    err = lib.ttx_last_rel_error()
    assert 0 <= err < 1e-4
    ptrs = ctypes.POINTER(ctypes.c_void_p)()
    used = (ctypes.c_int32 * 17)()
    d = lib.ttx_last_cores(ctypes.byref(ptrs), used)
    assert d == 6 and list(used[:7]) == ranks
    # the side-channel cores reproduce the tensor (device -> host through the HIP runtime)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hosts = []
    for k in range(6):
        if kind == "tr" and k == 0:
            n = dims[0] * ranks[0] * ranks[1]
        elif kind == "tr" and k == 5:
            n = ranks[5] * dims[5] * ranks[0]
        else:
            n = ranks[k] * dims[k] * ranks[k + 1]
        h = np.empty(n, dtype=np.float32)
        assert hip.hipMemcpy(h.ctypes.data, ptrs[k], 4 * n, 2) == 0     # hipMemcpyDeviceToHost
        hosts.append(h)
    An = before.astype(np.float64)
    full = ttsvd.tr_full(hosts, dims, ranks) if kind == "tr" else ttsvd.tt_full(hosts, dims, ranks)
    assert ttsvd.rel_error(An, full) < 1e-4


THIRD = {"tt": "_Z15ttdec_gpu_floatPfPiS0_PdS_S_", "tr": "_Z15trdec_gpu_floatPfPiS0_PdS_S_"}


def _load3(kind):
    lib = _load(kind)
    fn = getattr(lib, THIRD[kind])
    fn.restype = ctypes.c_float
    PI, PF = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)
    fn.argtypes = [PF, PI, PI, ctypes.POINTER(ctypes.c_double), PF, PF]
    return lib, fn


@pytest.mark.parametrize("kind", ["tt", "tr"])
def test_third_order_symbols_demangle_to_reference_signatures(kind):
    """ttdec_gpu_float / trdec_gpu_float (third_order_GPUs/opt/func.h:16, opt_GPU/func.h:15)."""
    lib, _ = _load3(kind)
    cxxfilt = shutil.which("c++filt")
    if cxxfilt is None:
        pytest.skip("c++filt not available")
    out = subprocess.run([cxxfilt, THIRD[kind]], stdout=subprocess.PIPE, text=True).stdout.strip()
    name = "ttdec_gpu_float" if kind == "tt" else "trdec_gpu_float"
    assert out == name + "(float*, int*, int*, double*, float*, float*)"


def _cores_to_host(lib, dims, kind):
    ptrs = ctypes.POINTER(ctypes.c_void_p)()
    used = (ctypes.c_int32 * 17)()
    d = lib.ttx_last_cores(ctypes.byref(ptrs), used)
    assert d == len(dims)
    ranks = list(used[:d + 1])
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hosts = []
    for k in range(d):
        n = ranks[k] * dims[k] * ranks[k + 1]
        h = np.empty(n, dtype=np.float32)
        assert hip.hipMemcpy(h.ctypes.data, ptrs[k], 4 * n, 2) == 0     # hipMemcpyDeviceToHost
        hosts.append(h)
    return hosts, ranks


@pytest.mark.gpu
@pytest.mark.parametrize("kind,gen_ranks,ranks", [
    ("tt", [1, 5, 4, 1], [1, 5, 4, 1]),      # exact ranks: error at the noise level
    ("tt", [1, 6, 5, 1], [1, 3, 3, 1]),      # truncated: an error of order one
    ("tr", [2, 2, 2, 2], [2, 2, 8, 2]),      # ring of ranks 2, inner rank r0^2 r2 reproduces it
    ("tr", [2, 3, 3, 2], [2, 2, 4, 2]),      # truncated ring
])
def test_third_order_driver_fills_error_and_ratio(engine, kind, gen_ranks, ranks):
    """The third-order drivers run like the reference's main (one decomposition, seconds back) and
    fill singleError / singleComRatio, which the reference declares and leaves unset: the error
    agrees with the oracle's reconstruction from the returned cores, and the ratio is the tensor's
    entries over the cores' entries."""
    lib, fn = _load3(kind)
    dims = [24, 20, 16]
    A = engine.generate(dims, gen_ranks, std=1.0, noise_rel=1e-6, seed=1134, tr=(kind == "tr"))
    hA = np.ascontiguousarray(A.cpu().numpy(), dtype=np.float32)
    before = hA.copy()
    n3 = (ctypes.c_int * 3)(*dims)
    r4 = (ctypes.c_int * 4)(*ranks)
    eps = (ctypes.c_double * 2)(1e-6, 1e-8)
    err = ctypes.c_float(-7.0)
    ratio = ctypes.c_float(-7.0)
    t = fn(hA.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n3, r4, eps, ctypes.byref(err), ctypes.byref(ratio))
    assert t > 0
    assert np.array_equal(hA, before)
    hosts, used = _cores_to_host(lib, dims, kind)
    assert used == ranks
    full = ttsvd.tr_full(hosts, dims, used) if kind == "tr" else ttsvd.tt_full(hosts, dims, used)
    e_or = ttsvd.rel_error(before.astype(np.float64), full)
    assert abs(err.value - e_or) <= 2e-7 + 1e-4 * e_or
    if ranks == gen_ranks or (kind == "tr" and ranks[2] >= ranks[0] ** 2 * gen_ranks[2]):
        assert err.value < 1e-4
    else:
        assert err.value > 1e-3
    core_entries = sum(used[k] * dims[k] * used[k + 1] for k in range(3))
    assert ratio.value == pytest.approx(np.prod(dims) / core_entries, rel=1e-6)
    # a null singleError / singleComRatio is allowed (the reference's callers may pass either)
    t2 = fn(hA.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n3, r4, eps, None, None)
    assert t2 > 0
