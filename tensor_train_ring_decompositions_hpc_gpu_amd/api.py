"""Python mirror of the reference's TT-SVD / TR-SVD entry points over the ttx C ABI.

PyTorch is used only for device memory and the stream: every tensor argument is a contiguous
float32 CUDA tensor holding column-major data (element (i1..id) at i1 + N1 (i2 + N2 (...)));
results come back as flat tensors together with their column-major shapes.  All arithmetic runs in
``libttx.so``; a missing library raises instead of falling back.

Reference entry points mirrored here:
* ``tt_svd``  -- ttdec_half_lu_device_6, ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:799-958
* ``tr_svd``  -- ttdec_half_lu_device_6 (TR), trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:827-1004
* ``rsvd``    -- basicRandSVD_LU, ttdecomposition/third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:557-775
* ``permute_231`` -- tensorToMode231, trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:49-58
* ``tt_rel_error`` / ``tr_rel_error`` -- calMSE (non-destructive), ttdec6_singleGPU.cu:134-182,
  trdec6_singleGPU.cu:175-228
"""
import ctypes

import numpy as np

from . import _lib as L


def _torch():
    import torch
    return torch


def _check(lib, h, status):
    if status not in (L.OK, L.RESIDUAL_FALLBACK_USED):
        msg = lib.ttx_last_error(h)
        raise L.TtxError(status, msg.decode() if msg else "")
    return status


def _dev_ptr(t, flat=True):
    """Device pointer of a contiguous float32 CUDA tensor.  flat: the buffer is a column-major tensor
    or matrix, passed as a 1-d tensor (a C-contiguous N-d tensor would be read transposed); the rank
    model's row-major batches (flat=False) may be 2-d."""
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("ttx: expected a CUDA tensor")
    if t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError("ttx: expected a contiguous float32 tensor")
    if flat and t.dim() != 1:
        raise TypeError("ttx: expected a flat (1-d) tensor holding column-major data")
    return t.data_ptr()


def _host_ptr(t):
    torch = _torch()
    if isinstance(t, np.ndarray):
        # the engine reads the buffer as the column-major tensor: a flat array, or an N-d array in
        # Fortran order (a C-ordered N-d array would be read transposed)
        if t.dtype != np.float32 or not (t.ndim == 1 and t.flags["C_CONTIGUOUS"] or t.flags["F_CONTIGUOUS"]):
            raise TypeError("ttx: expected a float32 array, flat or in Fortran (column-major) order")
        return t.ctypes.data
    if not isinstance(t, torch.Tensor) or t.is_cuda:
        raise TypeError("ttx: expected a host (CPU) tensor")
    if t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError("ttx: expected a contiguous float32 tensor")
    if t.dim() != 1:
        raise TypeError("ttx: expected a flat (1-d) tensor holding column-major data")
    return t.data_ptr()


def save_cores(path, cores, dims, ranks, tr=False):
    """Write a decomposition to one .npz file (numpy's container; plain arrays only, nothing pickled):
    core_k is core k flattened in the engine's layout (TT: (r_{k-1} N_k) x r_k column-major; TR: the
    first core N_1 x r_0 r_1, the last r_{d-1} x N_d r_0), with dims, ranks and kind ("tt" / "tr").
    cores: as tt_svd / tr_svd return them ((tensor, shape) pairs), or tensors / arrays."""
    d = len(dims)
    if len(cores) != d or len(ranks) != d + 1:
        raise ValueError("ttx: need d cores and d + 1 ranks")
    arrays = {}
    for k, c in enumerate(cores):
        t = c[0] if isinstance(c, tuple) else c
        a = t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
        size = int(ranks[k]) * int(dims[k]) * int(ranks[k + 1])
        a = a.reshape(-1)
        if a.size < size:
            raise ValueError("ttx: core %d holds %d values, the ranks need %d" % (k, a.size, size))
        arrays["core_%d" % k] = np.ascontiguousarray(a[:size], dtype=np.float32)
    np.savez(path, dims=np.asarray(dims, dtype=np.int64), ranks=np.asarray(ranks, dtype=np.int32),
             kind=np.array("tr" if tr else "tt"), **arrays)


def load_cores(path, device="cuda"):
    """Read a file written by save_cores: (cores, dims, ranks, tr).  device=None keeps numpy arrays;
    otherwise the cores are float32 torch tensors on that device, ready for Engine.full / rel_error."""
    with np.load(path, allow_pickle=False) as z:
        dims = [int(x) for x in z["dims"]]
        ranks = [int(x) for x in z["ranks"]]
        tr = str(z["kind"]) == "tr"
        cores = [np.array(z["core_%d" % k], dtype=np.float32) for k in range(len(dims))]
    if device is not None:
        torch = _torch()
        cores = [torch.from_numpy(c).to(device) for c in cores]
    return cores, dims, ranks, tr


class Engine:
    """One ttx handle bound to the current torch stream of the current device."""

    def __init__(self, stream=None):
        torch = _torch()
        self.lib = L.load()
        if stream is None:
            stream = torch.cuda.current_stream()
        self.stream = stream
        h = ctypes.c_void_p()
        st = self.lib.ttx_create(ctypes.byref(h), ctypes.c_void_p(stream.cuda_stream))
        if st != L.OK:
            raise L.TtxError(st, "ttx_create failed")
        self.h = h
        self.model_loaded = False

    # ---- rank model ----------------------------------------------------------------------------
    def load_rank_model(self, path=None):
        """Load trained rank-model weights (raw float32, MLP_PARAMS values; default: the packaged
        rank_mlp.f32)."""
        path = path or L.MODEL_PATH
        _check(self.lib, self.h, self.lib.ttx_mlp_load(self.h, path.encode()))
        self.model_loaded = True

    def mlp_init(self, seed=7):
        _check(self.lib, self.h, self.lib.ttx_mlp_init(self.h, seed))
        self.model_loaded = True

    def mlp_set_params(self, params):
        v = np.ascontiguousarray(params, dtype=np.float32)
        assert v.size == L.MLP_PARAMS
        _check(self.lib, self.h, self.lib.ttx_mlp_set_params(self.h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        self.model_loaded = True

    def mlp_get_params(self):
        v = np.zeros(L.MLP_PARAMS, dtype=np.float32)
        _check(self.lib, self.h, self.lib.ttx_mlp_get_params(self.h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return v

    def mlp_forward(self, X):
        torch = _torch()
        Y = torch.empty(X.shape[0] * L.MLP_OUT, device=X.device, dtype=torch.float32)
        _check(self.lib, self.h, self.lib.ttx_mlp_forward(self.h, ctypes.c_void_p(_dev_ptr(X, flat=False)), X.shape[0],
                                                          ctypes.c_void_p(Y.data_ptr())))
        return Y.view(X.shape[0], L.MLP_OUT)

    def mlp_gradient(self, X, T):
        torch = _torch()
        g = torch.empty(L.MLP_PARAMS, device=X.device, dtype=torch.float32)
        loss = ctypes.c_double(0.0)
        _check(self.lib, self.h, self.lib.ttx_mlp_gradient(self.h, ctypes.c_void_p(_dev_ptr(X, flat=False)),
                                                           ctypes.c_void_p(_dev_ptr(T, flat=False)), X.shape[0],
                                                           ctypes.c_void_p(g.data_ptr()), ctypes.byref(loss)))
        return g, float(loss.value)

    def mlp_train_step(self, X, T, lr, want_loss=False):
        loss = ctypes.c_double(0.0)
        _check(self.lib, self.h, self.lib.ttx_mlp_train_step(self.h, ctypes.c_void_p(_dev_ptr(X, flat=False)),
                                                             ctypes.c_void_p(_dev_ptr(T, flat=False)), X.shape[0], float(lr),
                                                             ctypes.byref(loss) if want_loss else None))
        return float(loss.value) if want_loss else None

    def mlp_features(self, pivots_dev, rel_tol, m, n):
        torch = _torch()
        x = torch.zeros(L.MLP_IN, device=pivots_dev.device, dtype=torch.float32)
        if pivots_dev.dtype != torch.float64 or not pivots_dev.is_contiguous():
            raise TypeError("ttx: pivots must be a contiguous float64 CUDA tensor")
        _check(self.lib, self.h, self.lib.ttx_mlp_features(self.h, ctypes.c_void_p(pivots_dev.data_ptr()),
                                                           pivots_dev.numel(), float(rel_tol), int(m), int(n),
                                                           ctypes.c_void_p(x.data_ptr())))
        return x

    def mlp_predict(self, pivots_dev, rel_tol, m, n):
        """features + forward of one sample in one launch (the prediction a decomposition step makes)"""
        torch = _torch()
        y = torch.zeros(L.MLP_OUT, device=pivots_dev.device, dtype=torch.float32)
        if pivots_dev.dtype != torch.float64 or not pivots_dev.is_contiguous():
            raise TypeError("ttx: pivots must be a contiguous float64 CUDA tensor")
        _check(self.lib, self.h, self.lib.ttx_mlp_predict(self.h, ctypes.c_void_p(pivots_dev.data_ptr()),
                                                          pivots_dev.numel(), float(rel_tol), int(m), int(n),
                                                          ctypes.c_void_p(y.data_ptr())))
        return y

    def probe(self, A, m, n, seed=1234, stream=0):
        piv = np.zeros(64, dtype=np.float64)
        _check(self.lib, self.h, self.lib.ttx_probe(self.h, ctypes.c_void_p(_dev_ptr(A)), int(m), int(n), seed, stream,
                                                    piv.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return piv

    def close(self):
        if self.h:
            self.lib.ttx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------------------------------
    def params(self, dims, ranks, s=10, q=2, seed=1234, precision="bf16", method="auto", rank_mode="fixed",
               eps=1e-4, flags=0):
        p = L.Params()
        d = len(dims)
        self.lib.ttx_params_init(ctypes.byref(p), d, L.i64_array(dims), L.i32_array(ranks))
        p.oversampling = int(s)
        p.power_iters = int(q)
        p.omega_seed = int(seed)
        p.precision = L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        p.method = L.METHODS[method] if isinstance(method, str) else int(method)
        p.rank_mode = L.RANK_MODES[rank_mode] if isinstance(rank_mode, str) else int(rank_mode)
        p.eps = float(eps)
        p.flags = int(flags)
        return p

    def _plan(self, dims, ranks, tr, kw):
        """Parameter block and core sizes of a call, kept per call signature: the entry points take the
        block as const, so repeated calls with the same arguments reuse it (a loop of decompositions
        then spends its host time in the engine, not in building ctypes structures)."""
        try:
            key = (tuple(dims), tuple(ranks), tr, tuple(kw.items()))
            hash(key)
        except TypeError:
            key = None
        cache = self.__dict__.setdefault("_plans", {})
        hit = cache.get(key) if key is not None else None
        if hit is not None:
            return hit
        p = self.params(dims, ranks, **kw)
        sizes = [int(self.lib.ttx_core_size(ctypes.byref(p), int(tr), k)) for k in range(len(dims))]
        if key is not None:
            if len(cache) >= 64:
                cache.clear()
            cache[key] = (p, sizes)
        return p, sizes

    def _sweep(self, fn, A, dims, ranks, tr, cores=None, host_slice=None, **kw):
        torch = _torch()
        p, sizes = self._plan(dims, ranks, tr, kw)
        if p.rank_mode == L.RANK_MODES["learned"] and not self.model_loaded:
            self.load_rank_model()
        d = len(dims)
        if cores is not None:
            # caller-owned output buffers (core_buffers), written in place
            if len(cores) != d or any(not (c.is_cuda and c.dtype == torch.float32 and c.is_contiguous()) or
                                      c.numel() < sz for c, sz in zip(cores, sizes)):
                raise ValueError("ttx: cores must be d contiguous float32 CUDA tensors of at least ttx_core_size floats")
        else:
            dev = "cuda" if host_slice is not None else A.device
            cores = [torch.empty(sz, device=dev, dtype=torch.float32) for sz in sizes]
        sigma = np.zeros((d - 1) * L.MAX_ELL, dtype=np.float32)
        rep = L.Report()
        sig_p = sigma.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        core_p = L.ptr_array([c.data_ptr() for c in cores])
        if host_slice is None:
            st = fn(self.h, ctypes.c_void_p(_dev_ptr(A)), ctypes.byref(p), core_p, sig_p, ctypes.byref(rep))
        else:
            st = fn(self.h, ctypes.c_void_p(_host_ptr(A)), ctypes.byref(p), int(host_slice), core_p, sig_p,
                    ctypes.byref(rep))
        _check(self.lib, self.h, st)
        used = rep.ranks[:d + 1]
        out = []
        for k in range(d):
            if tr and k == 0:
                shape = (dims[0], used[0] * used[1])
            elif tr and k == d - 1:
                shape = (used[d - 1], dims[d - 1] * used[0])
            else:
                shape = (used[k] * dims[k], used[k + 1])
            out.append((cores[k][: shape[0] * shape[1]], shape))
        info = report_dict(rep, d)
        return out, sigma.reshape(d - 1, L.MAX_ELL), info

    def core_buffers(self, dims, ranks, tr=False, **kw):
        """Output buffers for tt_svd / tr_svd(..., cores=...): one flat float32 CUDA tensor per core
        of ttx_core_size floats (the largest the ranks allow)."""
        torch = _torch()
        p = self.params(dims, ranks, **kw)
        return [torch.empty(int(self.lib.ttx_core_size(ctypes.byref(p), int(tr), k)), device="cuda",
                            dtype=torch.float32) for k in range(len(dims))]

    def tt_svd(self, A, dims, ranks, **kw):
        """TT-SVD of the column-major tensor A; returns ([(core, (rows, cols))], sigma, report).
        cores=core_buffers(...) (optional): write the cores there instead of new tensors."""
        return self._sweep(self.lib.ttx_tt_svd, A, dims, ranks, False, **kw)

    def tt_svd_noalloc(self, A, dims, ranks, **kw):
        """tt_svd without device allocation inside the call (ttx_tt_svd_noalloc): the engine's work
        buffers come from its static scratch array; pass cores=core_buffers(...) so that nothing is
        allocated here either.  Same results as tt_svd."""
        return self._sweep(self.lib.ttx_tt_svd_noalloc, A, dims, ranks, False, **kw)

    def tt_svd_host(self, A, dims, ranks, slice_cols=0, **kw):
        """TT-SVD of a column-major tensor held in host memory (a contiguous float32 CPU torch tensor,
        pinned for overlapped copies, or a numpy array): the first step streams it in column slices
        of slice_cols columns (0: about 64 MB each); the cores land on the GPU.  Returns as tt_svd."""
        return self._sweep(self.lib.ttx_tt_svd_host, A, dims, ranks, False, host_slice=slice_cols, **kw)

    def tr_svd_host(self, A, dims, ranks, slice_cols=0, **kw):
        """TR-SVD of a tensor in host memory (as tt_svd_host; first step of rank ranks[0] * ranks[1])."""
        return self._sweep(self.lib.ttx_tr_svd_host, A, dims, ranks, True, host_slice=slice_cols, **kw)

    def tr_svd(self, A, dims, ranks, **kw):
        """TR-SVD (first step rank ranks[0] * ranks[1]); returns ([(core, shape)], sigma, report)."""
        return self._sweep(self.lib.ttx_tr_svd, A, dims, ranks, True, **kw)

    def batch_tt_svd(self, tensors, dims_list, ranks_list, s=10, q=2, seed=1234, precision="bf16", flags=0,
                     streams=4, graph=True, autotune=True, per_tensor=False, cores=None, engine="grouped"):
        """TT-SVD of many tensors with fixed ranks (ttx_batch_tt_svd).  tensors: list of contiguous float32
        CUDA tensors; all of one order.  engine: "grouped" (one launch per kernel variant and step phase
        for the whole batch; FP16 batches run on worker streams) or "streams" (each tensor's sweep on one
        of ``streams`` worker streams).  Returns (cores per tensor [(core, shape)], report dict).  Pass the
        returned core buffers back as ``cores`` to relaunch the same captured graph."""
        if engine not in ("grouped", "streams"):
            raise ValueError("ttx: engine is 'grouped' or 'streams'")
        torch = _torch()
        count = len(tensors)
        d = len(dims_list[0])
        if any(len(x) != d for x in dims_list) or any(len(r) != d + 1 for r in ranks_list):
            raise ValueError("ttx: one order for the whole batch")
        if cores is None:
            cores = []
            for dims, ranks in zip(dims_list, ranks_list):
                p = self.params(dims, ranks)
                cores.append([torch.empty(int(self.lib.ttx_core_size(ctypes.byref(p), 0, k)), device="cuda",
                                          dtype=torch.float32) for k in range(d)])
        dims_flat = L.i64_array([x for dims in dims_list for x in dims])
        ranks_flat = L.i32_array([x for ranks in ranks_list for x in ranks])
        tptr = L.ptr_array([_dev_ptr(t) for t in tensors])
        cptr = L.ptr_array([c.data_ptr() for cs in cores for c in cs])
        common = self.params(dims_list[0], ranks_list[0], s=s, q=q, seed=seed, precision=precision, flags=flags)
        bflags = (L.BATCH_GRAPH if graph else 0) | (L.BATCH_AUTOTUNE if autotune else 0) | \
                 (L.BATCH_PER_TENSOR if per_tensor else 0) | (L.BATCH_STREAMS if engine == "streams" else 0)
        rep = L.BatchReport()
        st = self.lib.ttx_batch_tt_svd(self.h, count, d, dims_flat, ranks_flat, tptr, cptr, ctypes.byref(common),
                                       streams, bflags, ctypes.byref(rep))
        _check(self.lib, self.h, st)
        out = []
        for cs, dims, ranks in zip(cores, dims_list, ranks_list):
            shapes = [(ranks[k] * dims[k], ranks[k + 1]) for k in range(d - 1)] + [(ranks[d - 1], dims[d - 1])]
            out.append([(c[: a * b], (a, b)) for c, (a, b) in zip(cs, shapes)])
        info = {"total_ms": rep.total_ms, "tensors_per_s": rep.tensors_per_s, "p50_ms": rep.p50_ms,
                "p99_ms": rep.p99_ms, "streams": rep.streams, "graph": rep.graph, "graph_reused": rep.graph_reused,
                "tuned_shapes": rep.tuned_shapes, "engine": "grouped" if rep.engine == 1 else "streams",
                "launches": rep.launches, "sketch_chunks": rep.sketch_chunks, "core_buffers": cores}
        return out, info

    def rsvd(self, A, m, n, k, s=10, q=2, seed=1234, precision="bf16"):
        torch = _torch()
        ell = k + s
        U = torch.empty(m * ell, device=A.device, dtype=torch.float32)
        S = torch.empty(ell, device=A.device, dtype=torch.float32)
        SVT = torch.empty(ell * n, device=A.device, dtype=torch.float32)
        prec = L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        st = self.lib.ttx_rsvd(self.h, ctypes.c_void_p(_dev_ptr(A)), m, n, k, s, q, seed, prec,
                               ctypes.c_void_p(U.data_ptr()), ctypes.c_void_p(S.data_ptr()),
                               ctypes.c_void_p(SVT.data_ptr()))
        _check(self.lib, self.h, st)
        return U, S, SVT

    def step(self, A, m, n, r, s=10, q=2, seed=1234, stream_id=0, precision="bf16", method="auto"):
        torch = _torch()
        core = torch.empty(m * r, device=A.device, dtype=torch.float32)
        nxt = torch.empty(r * n, device=A.device, dtype=torch.float32)
        sigma = np.zeros(L.MAX_ELL, dtype=np.float32)
        path = ctypes.c_int32(0)
        prec = L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        meth = L.METHODS[method] if isinstance(method, str) else int(method)
        st = self.lib.ttx_step(self.h, ctypes.c_void_p(_dev_ptr(A)), m, n, r, s, q, seed, stream_id, prec, meth,
                               ctypes.c_void_p(core.data_ptr()), ctypes.c_void_p(nxt.data_ptr()),
                               sigma.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(path))
        _check(self.lib, self.h, st)
        return core, nxt, sigma, int(path.value)

    def permute_231(self, x, r0, r1, n):
        torch = _torch()
        out = torch.empty(r0 * r1 * n, device=x.device, dtype=torch.float32)
        st = self.lib.ttx_permute_231(self.h, ctypes.c_void_p(_dev_ptr(x)), ctypes.c_void_p(out.data_ptr()),
                                      r0, r1, n)
        _check(self.lib, self.h, st)
        return out

    def fill_normal(self, count, seed, stream, scale=1.0, device="cuda"):
        torch = _torch()
        out = torch.empty(count, device=device, dtype=torch.float32)
        _check(self.lib, self.h, self.lib.ttx_fill_normal(self.h, ctypes.c_void_p(out.data_ptr()), count, seed,
                                                          stream, scale))
        return out

    def omega(self, n, ell, seed, stream, device="cuda"):
        torch = _torch()
        out = torch.empty(n * ell, device=device, dtype=torch.float32)
        _check(self.lib, self.h, self.lib.ttx_omega(self.h, ctypes.c_void_p(out.data_ptr()), n, ell, seed, stream))
        return out

    def generate(self, dims, ranks, std=1.0, noise_rel=0.0, seed=1134, tr=False, return_cores=False,
                 device="cuda"):
        torch = _torch()
        d = len(dims)
        total = int(np.prod(dims))
        A = torch.empty(total, device=device, dtype=torch.float32)
        cores = None
        cptr = None
        if return_cores:
            p = self.params(dims, ranks)
            cores = [torch.empty(int(self.lib.ttx_core_size(ctypes.byref(p), int(tr), k)), device=device,
                                 dtype=torch.float32) for k in range(d)]
            cptr = L.ptr_array([c.data_ptr() for c in cores])
        fn = self.lib.ttx_tr_generate if tr else self.lib.ttx_tt_generate
        st = fn(self.h, d, L.i64_array(dims), L.i32_array(ranks), std, noise_rel, seed,
                ctypes.c_void_p(A.data_ptr()), cptr)
        _check(self.lib, self.h, st)
        return (A, cores) if return_cores else A

    def generate_ref(self, dims, ranks, seed=1134, tr=False, return_cores=False, device="cuda"):
        """The reference harness's input (generalTTTensor / generalTRTensor, ttdec6_singleGPU.cu:185-271,
        trdec6_singleGPU.cu:231-329): U[-1, 1] cores from glibc's rand() after srand(seed), drawn in the
        order G_{d-1}, G_d, G_{d-2}, ..., G_1, contracted on the device."""
        torch = _torch()
        d = len(dims)
        A = torch.empty(int(np.prod(dims)), device=device, dtype=torch.float32)
        cores, cptr = None, None
        if return_cores:
            p = self.params(dims, ranks)
            cores = [torch.empty(int(self.lib.ttx_core_size(ctypes.byref(p), int(tr), k)), device=device,
                                 dtype=torch.float32) for k in range(d)]
            cptr = L.ptr_array([c.data_ptr() for c in cores])
        fn = self.lib.ttx_tr_generate_ref if tr else self.lib.ttx_tt_generate_ref
        _check(self.lib, self.h, fn(self.h, d, L.i64_array(dims), L.i32_array(ranks), int(seed),
                                    ctypes.c_void_p(A.data_ptr()), cptr))
        return (A, cores) if return_cores else A

    def rel_error(self, cores, dims, ranks, A, tr=False):
        err = ctypes.c_double(0.0)
        fn = self.lib.ttx_tr_rel_error if tr else self.lib.ttx_tt_rel_error
        ptrs = L.ptr_array([(c[0] if isinstance(c, tuple) else c).data_ptr() for c in cores])
        st = fn(self.h, len(dims), L.i64_array(dims), L.i32_array(ranks), ptrs, ctypes.c_void_p(_dev_ptr(A)),
                ctypes.byref(err))
        _check(self.lib, self.h, st)
        return float(err.value)

    def tt_round(self, cores, dims, ranks, eps, max_ranks=None):
        """TT rounding of cores (ttx_tt_round): recompress to the smallest ranks with relative error at most
        eps (capped by max_ranks).  cores: device tensors in the TT layout (or (tensor, shape) pairs).
        Returns ([(core, (rows, cols))], new ranks, sigma (d-1) x TTX_MAX_ELL per bond)."""
        torch = _torch()
        d = len(dims)
        if len(cores) != d or len(ranks) != d + 1:
            raise ValueError("ttx: need d cores and d + 1 ranks")
        ins = [(c[0] if isinstance(c, tuple) else c) for c in cores]
        for k, c in enumerate(ins):
            _dev_ptr(c)
            if c.numel() < ranks[k] * dims[k] * ranks[k + 1]:
                raise ValueError("ttx: core %d holds fewer values than its ranks need" % k)
        outs = [torch.empty(ranks[k] * dims[k] * ranks[k + 1], device="cuda", dtype=torch.float32) for k in range(d)]
        rout = (ctypes.c_int32 * (d + 1))()
        sigma = np.zeros((d - 1) * L.MAX_ELL, dtype=np.float32)
        rmax = L.i32_array(max_ranks) if max_ranks is not None else None
        st = self.lib.ttx_tt_round(self.h, d, L.i64_array(dims), L.i32_array(ranks),
                                   L.ptr_array([c.data_ptr() for c in ins]), float(eps), rmax,
                                   L.ptr_array([c.data_ptr() for c in outs]), rout,
                                   sigma.ctypes.data_as(ctypes.c_void_p))
        _check(self.lib, self.h, st)
        rn = [int(x) for x in rout]
        shapes = [(rn[k] * dims[k], rn[k + 1]) for k in range(d)]
        return [(c[: a * b], (a, b)) for c, (a, b) in zip(outs, shapes)], rn, sigma.reshape(d - 1, L.MAX_ELL)

    def full(self, cores, dims, ranks, tr=False):
        torch = _torch()
        out = torch.empty(int(np.prod(dims)), device="cuda", dtype=torch.float32)
        fn = self.lib.ttx_tr_full if tr else self.lib.ttx_tt_full
        ptrs = L.ptr_array([(c[0] if isinstance(c, tuple) else c).data_ptr() for c in cores])
        st = fn(self.h, len(dims), L.i64_array(dims), L.i32_array(ranks), ptrs, ctypes.c_void_p(out.data_ptr()))
        _check(self.lib, self.h, st)
        return out


def reachable_ranks(dims, ranks):
    """TT ranks the sweep keeps: r_k = min(r_k, r_{k-1} N_k, N_{k+1} ... N_d), left to right"""
    out = [1]
    for k in range(1, len(dims)):
        m = out[k - 1] * dims[k - 1]
        n = int(np.prod(dims[k:]))
        out.append(int(min(ranks[k], m, n)))
    out.append(1)
    return out


def report_dict(rep, d):
    # ctypes array slices convert in one call (an indexed read per element costs a Python object each)
    n = d - 1
    return {
        "ranks": rep.ranks[:d + 1],
        "ell": rep.ell[:n],
        "path": rep.path[:n],
        "power_iters": rep.power_iters[:n],
        "zero_pivots": rep.zero_pivots[:n],
        "jacobi_sweeps": rep.jacobi_sweeps[:n],
        "norm2_in": rep.norm2_in[:n],
        "step_ms": rep.step_ms[:n],
        "total_ms": float(rep.total_ms),
        "bytes_alg": int(rep.bytes_alg),
        "range_resid": rep.range_resid[:n],
        "range_checks": rep.range_checks[:n],
        "passes": rep.passes[:n],
        "fallback": rep.fallback[:n],
        "predicted_rank": rep.predicted_rank[:n],
        "kernels": _kernel_classes(rep),
    }


def _kernel_classes(rep):
    launches = rep.kernel_launches[:]
    if not any(launches):
        return {}
    return {L.KCLASS[c]: {"ms": float(rep.kernel_ms[c]), "launches": int(launches[c]),
                          "bytes": int(rep.kernel_bytes[c]), "top_bytes": int(rep.kernel_top_bytes[c]),
                          "top_ms": float(rep.kernel_top_ms[c])}
            for c in range(L.NUM_KCLASS) if launches[c] > 0}
