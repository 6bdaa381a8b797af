"""The reference harness's own inputs and shapes.

generalTTTensor / generalTRTensor (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:185-271,
trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:231-329) build U[-1, 1] cores from glibc's rand()
after srand(1134); the harness mains (ttdec6_singleGPU.cu:962-1006, trdec6_singleGPU.cu:1008-1052)
decompose k^6 tensors: TT for k = 8, 12, ..., 44 with ranks 4 (k <= 24) or 8, the device driver below
k = 40 and the host driver from 40; TR for k = 40, 44, 48 with ranks {2, 4, 8, 8, 8, 8, 2} through the
host driver.  Here the engine's generator is checked bit for bit against the oracle's restatement of
glibc rand() (itself pinned to the C library), the small shapes against the FP64 oracle, and every
harness shape runs through the reference-name shims as the mains call them.

Sizes: 44^6 is 7.3e9 elements (29 GB), 48^6 12.2e9 (49 GB); the host buffers are numpy arrays, the
device holds the tensor, its host-streamed copy and the error's copy (about 150 GB at 48^6)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import glibc_rand, ttsvd

pytestmark = pytest.mark.gpu

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "tensor_train_ring_decompositions_hpc_gpu_amd")
DEVICE6 = "_Z22ttdec_half_lu_device_6PfPKiS1_RS0_S2_b"
HOST6 = "_Z20ttdec_half_lu_host_6PfPiPKiRS1_S3_b"
GEN = {"tt": "_Z15generalTTTensorPfPKiS1_RS0_", "tr": "_Z15generalTRTensorPfPKiS1_RS0_"}
LIBS = {"tt": "libttx_ttdec6.so", "tr": "libttx_trdec6.so"}


def _shapes(dims, ranks, tr):
    d = len(dims)
    if not tr:
        return [(ranks[k] * dims[k], ranks[k + 1]) for k in range(d)]
    return ([(dims[0], ranks[0] * ranks[1])] + [(ranks[k] * dims[k], ranks[k + 1]) for k in range(1, d - 1)] +
            [(ranks[d - 1], dims[d - 1] * ranks[0])])


@pytest.mark.parametrize("tr", [False, True])
def test_generate_ref_bit_exact(engine, tr):
    """Cores bit-identical to the oracle's glibc restatement (reference order G5, G6, G4, G3, G2, G1);
    the contracted tensor equals the FP64 contraction of those cores to FP32 rounding."""
    dims = [5, 6, 7, 4, 6, 5]
    ranks = [2, 3, 4, 2, 3, 2, 2] if tr else [1, 3, 4, 2, 3, 2, 1]
    A, cores = engine.generate_ref(dims, ranks, tr=tr, return_cores=True)
    ref = glibc_rand.ref_cores(dims, ranks, tr=tr, seed=1134)
    for c, r in zip(cores, ref):
        assert np.array_equal(c.cpu().numpy().view(np.uint32), r.view(np.uint32))
    cs = [r.reshape(a, b, order="F") for r, (a, b) in zip(ref, _shapes(dims, ranks, tr))]
    full = ttsvd.tr_full(cs, dims, ranks) if tr else ttsvd.tt_full(cs, dims, ranks)
    got = A.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(got - full)) <= 1e-5 * np.max(np.abs(full))


def _lib(kind):
    path = os.path.join(PKG, LIBS[kind])
    if not os.path.exists(path):
        pytest.skip("shim library not built")
    lib = ctypes.CDLL(path)
    PI = ctypes.POINTER(ctypes.c_int)
    for name in (DEVICE6, HOST6):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_float
        fn.argtypes = [ctypes.POINTER(ctypes.c_float), PI, PI, PI, PI, ctypes.c_bool]
    gen = getattr(lib, GEN[kind])
    gen.restype = None
    gen.argtypes = [ctypes.POINTER(ctypes.c_float), PI, PI, PI]
    lib.ttx_last_cores.restype = ctypes.c_int32
    lib.ttx_last_cores.argtypes = [ctypes.POINTER(ctypes.POINTER(ctypes.c_void_p)), ctypes.POINTER(ctypes.c_int32)]
    lib.ttx_last_rel_error.restype = ctypes.c_double
    return lib


def _harness(kind, k, ranks, name, store=None):
    """The harness main's calls for one k: generate into the host buffer (the first k^6 floats of
    store, or a new array), decompose with calError, read the error.  Returns (seconds, error, ranks
    used, the host tensor)."""
    lib = _lib(kind)
    dims = [k] * 6
    total = k ** 6
    buf = store[:total] if store is not None else np.empty(total, dtype=np.float32)
    r = (ctypes.c_int * 7)(*ranks)
    d = (ctypes.c_int * 6)(*dims)
    pf = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    getattr(lib, GEN[kind])(pf, r, d, ctypes.byref(ctypes.c_int(64)))
    t = getattr(lib, name)(pf, r, d, ctypes.byref(ctypes.c_int(1)), ctypes.byref(ctypes.c_int(64)), True)
    used = (ctypes.c_int32 * 17)()
    order = lib.ttx_last_cores(ctypes.byref(ctypes.POINTER(ctypes.c_void_p)()), used)
    return t, lib.ttx_last_rel_error(), list(used[:7]) if order == 6 else None, buf


@pytest.mark.parametrize("kind,k,ranks", [
    ("tt", 8, [1, 4, 4, 4, 4, 4, 1]),
    ("tt", 12, [1, 4, 4, 4, 4, 4, 1]),
    ("tr", 8, [2, 2, 4, 4, 4, 4, 2]),
    ("tr", 12, [2, 2, 4, 4, 4, 4, 2]),
])
def test_harness_small_shapes_match_oracle(engine, monkeypatch, kind, k, ranks):
    """The harness's first shapes in FP32 (TTX_PRECISION=fp32, s = 10, q = 2 fixed): the shim's
    reconstruction error is the FP64 oracle's on the same generated tensor (TT: exact ranks, error at
    FP32 rounding; TR: the harness's ranks truncate the ring, error of order one)."""
    monkeypatch.setenv("TTX_PRECISION", "fp32")
    monkeypatch.setenv("TTX_ADAPTIVE", "0")
    t, err, used, hA = _harness(kind, k, ranks, DEVICE6)
    assert t > 0 and used == ranks
    A = hA.astype(np.float64)
    dims = [k] * 6
    if kind == "tt":
        c, _, _, u = ttsvd.tt_svd(A, dims, ranks, s=10, q=2)
        e_ref = ttsvd.rel_error(A, ttsvd.tt_full(c, dims, u))
        assert e_ref < 1e-7 and err < 2e-6, (err, e_ref)
    else:
        c, _, _, u = ttsvd.tr_svd(A, dims, ranks, s=10, q=2)
        e_ref = ttsvd.rel_error(A, ttsvd.tr_full(c, dims, u))
        assert abs(err - e_ref) <= 1e-5 + 1e-4 * e_ref, (err, e_ref)


def test_harness_tt_shapes():
    """Every shape of the TT harness main (ttdec6_singleGPU.cu:962-1006) through the shims as it calls
    them (the device driver for k < 40, the host driver from 40, p = 64 slices, calError), BF16 with
    the adaptive range check (the shims' defaults): exact-rank inputs reconstruct to the 1e-5 level."""
    store = np.empty(44 ** 6, dtype=np.float32)
    for k in range(8, 48, 4):
        r = 8 if k > 24 else 4
        ranks = [1, r, r, r, r, r, 1]
        t, err, used, _ = _harness("tt", k, ranks, DEVICE6 if k < 40 else HOST6, store)
        print("k=%d ranks %d: %.4f s, rel error %.3e" % (k, r, t, err), flush=True)
        assert t > 0 and used == ranks, k
        assert 0 <= err < 1e-5, (k, err)


def test_harness_tr_shapes(engine):
    """Every shape of the TR harness main (trdec6_singleGPU.cu:1008-1052: k = 40, 44, 48, ranks
    {2, 4, 8, 8, 8, 8, 2}, the host driver with p = 64 slices): the harness's ranks truncate the ring
    (the oracle's error at k = 8, 12 with these ranks is 0.735, 0.886), so the check is that the
    host-streamed driver's error equals the device-resident sweep's on the same tensor (1e-3)."""
    import torch
    ranks = [2, 4, 8, 8, 8, 8, 2]
    store = np.empty(48 ** 6, dtype=np.float32)
    for k in (40, 44, 48):
        t, err, used, hA = _harness("tr", k, ranks, HOST6, store)
        assert t > 0 and used == ranks, k
        dims = [k] * 6
        dA = torch.from_numpy(hA).cuda()
        cores, _, rep = engine.tr_svd(dA, dims, ranks, s=10, q=2, flags=4)
        e_dev = engine.rel_error(cores, dims, ranks, dA, tr=True)
        print("k=%d: %.4f s, rel error host %.6f device %.6f" % (k, t, err, e_dev), flush=True)
        # the same Omega streams in both drivers; only the first step's Gram is summed in another order
        assert 0.5 < err < 1.0 and abs(err - e_dev) <= 1e-3 * e_dev, (k, err, e_dev)
        del dA, cores, hA
        torch.cuda.empty_cache()
