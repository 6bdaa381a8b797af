"""Throughput of the TT-SVD hot path on one MI355X.

Workload (one "step" = one full TT-SVD of one tensor already resident in HBM): order-6 tensor
32^6 (1.07e9 FP32 elements, 4.295 GB) built from seeded TT cores with ranks (1,32,32,32,32,32,1),
entries N(0, (1/sqrt(32))^2), plus 1e-6 relative Gaussian noise; sketch width ell = r + s with
s = 10.  The headline engine runs BF16 MFMA (hi/lo split, FP32 accumulation); the "FP32 plain"
comparison runs the same sweep with FP32 MFMA arithmetic and q = 2 power iterations.

value = input GB/s = 4.295e9 bytes / (seconds per decomposition).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TT-SVD input GB/s & %HBM peak, BF16+learned-rank vs FP32 plain at err≤1e-4"
HBM_PEAK_GBS = 8000.0
DIMS = [32] * 6
RANKS = [1, 32, 32, 32, 32, 32, 1]
DATA_SEED = 1134
OMEGA_SEED = 1234


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--q", type=int, default=2, help="power iterations of the headline run")
    ap.add_argument("--s", type=int, default=10, help="oversampling")
    ap.add_argument("--fast-range", action="store_true", help="single bf16 MFMA in the range finder")
    ap.add_argument("--tail-x4", action="store_true",
                    help="four bf16 products in the randomized steps' projection (default: three)")
    ap.add_argument("--no-adaptive", action="store_true",
                    help="headline leg with a fixed q instead of the a-posteriori range check")
    ap.add_argument("--rank-mode", default="learned", choices=["learned", "fixed"],
                    help="headline leg: learned ranks (eps 1e-4, ranks 32 as ceilings) or the fixed ranks 32")
    ap.add_argument("--only", default="", choices=["", "bf16", "fp32", "tr", "batch", "learned", "cfg1"],
                    help="run one leg (profiling); tr / batch / learned / cfg1 print that leg's JSON only")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-rank workload")
    ap.add_argument("--no-tr", action="store_true", help="skip the TR-SVD workload")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-host", action="store_true", help="skip the host-streamed workload")
    ap.add_argument("--batch-grouped-only", action="store_true",
                    help="batched workload: the grouped engine's runs only (no per-tensor or worker-stream timing)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 1 GiB device-to-device copy (scale reference for rocprof HBM counters)")
    return ap.parse_args()


def run_leg(eng, A, torch, prec, q, s, steps, warmup, flags=0, eps=1e-4, rank_mode="fixed"):
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    kw = dict(s=s, q=q, seed=OMEGA_SEED, precision=prec, flags=flags, eps=eps, rank_mode=rank_mode)
    # output buffers allocated once, outside the timed region (the library writes the cores there)
    bufs = eng.core_buffers(DIMS, RANKS, **kw)
    for _ in range(warmup):
        eng.tt_svd(A, DIMS, RANKS, cores=bufs, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        cores, sig, rep = eng.tt_svd(A, DIMS, RANKS, cores=bufs, **kw)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    # one extra instrumented run for the per-kernel device times (not part of the timed region)
    cores, sig, rep_t = eng.tt_svd(A, DIMS, RANKS, **dict(kw, flags=flags | L.FLAG_TIMING))
    err = eng.rel_error(cores, DIMS, rep_t["ranks"], A)
    return t, err, rep_t, sig


def learned_leg(eng, torch, L, count=4, dims=(48, 48, 48, 48, 48), eps=1e-3, eval_count=32):
    """Learned-rank workload: order-5 48^5 tensors (1.02 GB) whose bonds carry weights i^-alpha,
    alpha ~ U[0.5, 3] per bond (seeded), eps = 1e-3.  The same sweep with the trained network
    (LEARNED) and without it (DELTA: the adaptive range finder climbs 16, 32, ... widths)."""
    from tensor_train_ring_decompositions_hpc_gpu_amd import learned as Lr
    if not os.path.exists(L.MODEL_PATH):
        return None
    eng.load_rank_model()
    dims = list(dims)
    res = {m: {"ms": 0.0, "bytes": 0, "passes": 0} for m in ("delta", "learned")}
    hits = near = fallbacks = preds = 0
    abs_err = 0
    errs = []
    for seed in range(count):
        alphas = np.random.default_rng(seed).uniform(0.5, 3.0, len(dims) - 1)
        A, r = Lr.spectral_tt(dims, 48, alphas, seed)
        dA = torch.from_numpy(A).cuda()
        del A
        ceil = [1] + [min(48, x) for x in r[1:-1]] + [1]
        reps = {}
        for mode in ("delta", "learned"):
            kw = dict(s=8, q=2, precision="bf16", rank_mode=mode, eps=eps)
            eng.tt_svd(dA, dims, ceil, **kw)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                cores, sig, rep = eng.tt_svd(dA, dims, ceil, **kw)
            torch.cuda.synchronize()
            res[mode]["ms"] += (time.perf_counter() - t0) / 3 * 1e3
            res[mode]["bytes"] += rep["bytes_alg"]
            res[mode]["passes"] += sum(rep["passes"])
            errs.append(eng.rel_error(cores, dims, rep["ranks"], dA))
            reps[mode] = rep
        for k, pr in enumerate(reps["learned"]["predicted_rank"]):
            if pr > 0:
                preds += 1
                ref = reps["delta"]["ranks"][k + 1]
                hits += int(pr == ref)
                near += int(abs(pr - ref) <= 2)
                abs_err += abs(pr - ref)
                fallbacks += reps["learned"]["fallback"][k]
        del dA
    # the network's rank predictions over more tensors (order 5, 24^5, the same bond-weight law)
    ev = {"predictions": 0, "hits": 0, "near": 0, "abs": 0, "fallbacks": 0, "max_err": 0.0}
    for seed in range(100, 100 + eval_count):
        alphas = np.random.default_rng(seed).uniform(0.5, 3.0, 4)
        A, r = Lr.spectral_tt([24] * 5, 48, alphas, seed)
        dA = torch.from_numpy(A).cuda()
        ceil = [1] + [min(48, x) for x in r[1:-1]] + [1]
        reps = {}
        for mode in ("delta", "learned"):
            cores, sig, rep = eng.tt_svd(dA, [24] * 5, ceil, s=8, q=2, precision="bf16", rank_mode=mode, eps=eps)
            reps[mode] = rep
            ev["max_err"] = max(ev["max_err"], eng.rel_error(cores, [24] * 5, rep["ranks"], dA))
        for k, pr in enumerate(reps["learned"]["predicted_rank"]):
            if pr > 0:
                ref = reps["delta"]["ranks"][k + 1]
                ev["predictions"] += 1
                ev["hits"] += int(pr == ref)
                ev["near"] += int(abs(pr - ref) <= 2)
                ev["abs"] += abs(pr - ref)
                ev["fallbacks"] += reps["learned"]["fallback"][k]
        del dA
    npred = max(ev["predictions"], 1)
    evaluation = {"workload": "%d order-5 tensors 24^5, same bond-weight law, eps=1e-3, BF16, s=8, q<=2" % eval_count,
                  "predictions": ev["predictions"], "rank_hit_rate": round(ev["hits"] / npred, 3),
                  "rank_within_2_rate": round(ev["near"] / npred, 3), "rank_mae": round(ev["abs"] / npred, 3),
                  "fallback_rate": round(ev["fallbacks"] / npred, 3), "max_rel_error": ev["max_err"]}
    return {"workload": "%d order-5 tensors 48^5 (1.02 GB), bond weights i^-alpha, alpha~U[0.5,3], eps=1e-3, "
                        "BF16, s=8, q<=2" % count, "rank_model_evaluation": evaluation,
            "learned_ms": round(res["learned"]["ms"] / count, 3), "adaptive_only_ms": round(res["delta"]["ms"] / count, 3),
            "speedup": round(res["delta"]["ms"] / max(res["learned"]["ms"], 1e-9), 3),
            "bytes_learned": res["learned"]["bytes"], "bytes_adaptive_only": res["delta"]["bytes"],
            "passes_learned": res["learned"]["passes"], "passes_adaptive_only": res["delta"]["passes"],
            "rank_hit_rate": round(hits / max(preds, 1), 3), "rank_within_2_rate": round(near / max(preds, 1), 3),
            "rank_mae": round(abs_err / max(preds, 1), 3), "fallback_rate": round(fallbacks / max(preds, 1), 3),
            "predictions": preds, "max_rel_error": max(errs)}


def tr_leg(eng, torch, L, steps=3):
    """TR workload: order-6 32^6 tensor (4.295 GB) from ring cores 16 x 32 x 16 with N(0, 1/4) entries;
    TR-SVD with first-step split r0 r1 = 4 x 16 (64 columns on a 32-row unfolding: the core is zero-padded),
    inner ranks 16, FP16 range finder with power-of-two operand scaling, s = 10, q <= 2 (adaptive).
    The 231 relayout moves 2^31 elements."""
    dims = [32] * 6
    A = eng.generate(dims, [16] * 7, std=0.5, noise_rel=0.0, seed=DATA_SEED, tr=True)
    ranks = [4, 16, 16, 16, 16, 16, 4]
    kw = dict(s=10, q=2, seed=OMEGA_SEED, precision="fp16", flags=L.FLAG_ADAPTIVE)
    eng.tr_svd(A, dims, ranks, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        cores, sig, rep = eng.tr_svd(A, dims, ranks, **kw)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    cores, sig, rep = eng.tr_svd(A, dims, ranks, **dict(kw, flags=kw["flags"] | L.FLAG_TIMING))
    err = eng.rel_error(cores, dims, ranks, A, tr=True)
    c32, _, _ = eng.tr_svd(A, dims, ranks, **dict(kw, precision="fp32"))
    err32 = eng.rel_error(c32, dims, ranks, A, tr=True)
    ks = rep["kernels"]
    stream = {k: v for k, v in ks.items() if v["bytes"] > 0}
    dom = max(stream, key=lambda k: stream[k]["ms"])
    kd = ks[dom]
    achieved = kd["top_bytes"] / (kd["top_ms"] * 1e-3) / 1e9
    out = {"workload": "order-6 TR-SVD 32^6 (4.295 GB), ring cores 16x32x16 N(0,1/4), split r0 r1 = 4 x 16, "
                       "inner ranks 16, FP16 range finder, s=10, q<=2",
           "value": round(A.numel() * 4 / t / 1e9, 2), "unit": "GB/s", "ms_per_step": round(t * 1e3, 3),
           "rel_error_fp16": err, "rel_error_fp32_same_schedule": err32,
           "rel_error_note": "the split keeps a ring rank of 4 where the data has 16: the FP64 oracle's error for "
                             "this schedule is 0.842 / 0.932 / 0.978 / 0.987 at 6^6 / 8^6 / 12^6 / 16^6 "
                             "(tests/test_gpu_parity.py::test_tr_svd_reduced_cfg3_schedule checks 12^6)",
           "paths": rep["path"], "power_iters_run": rep["power_iters"], "passes": rep["passes"],
           "relayout": "231 permutation fused into the first step's projection store; rows beyond the first "
                       "step's rank bound (32 of 64) neither written nor read by step 2",
           "bytes_alg": rep["bytes_alg"], "pipeline_hbm_frac": round(rep["bytes_alg"] / (HBM_PEAK_GBS * 1e9) / t, 4),
           "baseline_roofline_frac": round(8.83e-3 / t, 4),
           "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": kd["top_bytes"],
                        "ms_per_launch": round(kd["top_ms"], 4)},
           "kernels": {k: {"ms": round(v["ms"], 4), "launches": v["launches"],
                           "GBs": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in ks.items()}}
    del A, cores, c32
    return out


def host_leg(eng, A, torch, L, steps=2):
    """Host-streamed workload: the same 32^6 tensor (4.295 GB) in pinned host memory; the first step
    reads it over PCIe in 64 column slices (64 MB each), the copies overlapped with the Gram kernels,
    once into a device copy (the default when it fits) or, with FLAG_STREAM_TWICE, twice through two
    slice buffers; the rest of the sweep as the fixed-rank BF16 leg (ranks 32, s = 10, q <= 2
    adaptive).  Also times one plain 4.295 GB host-to-device copy, the bound of one read."""
    Ah = torch.empty(A.numel(), dtype=torch.float32, pin_memory=True)
    Ah.copy_(A)
    kw = dict(s=10, q=2, seed=OMEGA_SEED, precision="bf16", flags=L.FLAG_ADAPTIVE)
    n = A.numel() // DIMS[0]
    slice_cols = n // 64
    bufs = eng.core_buffers(DIMS, RANKS, **kw)
    eng.tt_svd_host(Ah, DIMS, RANKS, slice_cols=slice_cols, cores=bufs, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        cores, sig, rep = eng.tt_svd_host(Ah, DIMS, RANKS, slice_cols=slice_cols, cores=bufs, **kw)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    err = eng.rel_error(cores, DIMS, rep["ranks"], A)
    # the bounded-memory schedule: two reads through two 64 MB slice buffers
    kw2 = dict(kw, flags=kw["flags"] | L.FLAG_STREAM_TWICE)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.tt_svd_host(Ah, DIMS, RANKS, slice_cols=slice_cols, cores=bufs, **kw2)
    torch.cuda.synchronize()
    t2 = time.perf_counter() - t0
    D = torch.empty_like(A)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    D.copy_(Ah.view_as(A), non_blocking=True)
    torch.cuda.synchronize()
    tc = time.perf_counter() - t0
    del D, Ah
    gb = A.numel() * 4 / 1e9
    return {"workload": "order-6 TT-SVD 32^6 (4.295 GB) from pinned host memory, 64 column slices, "
                        "fixed ranks 32, BF16, s=10, q<=2 (adaptive)",
            "value": round(gb / t, 2), "unit": "GB/s (tensor bytes / time, PCIe included)",
            "ms_per_step": round(t * 1e3, 3), "rel_error": err, "ranks_used": rep["ranks"],
            "h2d_copy_GBs": round(gb / tc, 2), "one_copy_ms": round(tc * 1e3, 3),
            "stream_twice": {"ms_per_step": round(t2 * 1e3, 3), "value": round(gb / t2, 2),
                             "two_copies_ms": round(2 * tc * 1e3, 3)}}


def batch_bytes_alg(dims_l, ranks_l, s=10, q=2):
    """Algorithmic HBM bytes of a batch: per tensor and step, P passes over the unfolding A_k plus the
    write of A_{k+1}, with P = 2 on an exact step and 2 q + 2 on a randomized one.  A batch step is exact
    when ell = r + s >= min(m, n) or min(m, n) <= 64 (the batch plan's rule)."""
    total = 0
    for dims, ranks in zip(dims_l, ranks_l):
        n = int(np.prod(dims))
        r_prev = 1
        for k in range(len(dims) - 1):
            m = r_prev * dims[k]
            n //= dims[k]
            r = min(ranks[k + 1], m, n)
            passes = 2 if (r + s >= min(m, n) or min(m, n) <= 64) else 2 * q + 2
            total += passes * m * n * 4 + r * n * 4
            r_prev = r
    return total


def batch_leg(eng, torch, count=512, streams=8, reps=5, grouped_only=False):
    """Batched workload: 512 order-4 tensors, each mode size uniform in {16..64} and each TT rank uniform
    in {4..32} per tensor (seeded), N(0,1) cores, 1e-5 relative noise; BF16 range finder, s = 10, q = 2.
    Grouped engine: every step phase of the whole batch is one launch per kernel variant, the launches
    captured once into a hipGraph and relaunched (median of reps launches), the sketch's chunks per
    workgroup autotuned for the batch; each tensor's own sweep (same plan) is timed alone for p50 / p99;
    the worker-stream engine (8 streams, eager, same plan) is timed beside it."""
    from tensor_train_ring_decompositions_hpc_gpu_amd.api import reachable_ranks
    rng = np.random.default_rng(DATA_SEED)
    tensors, dims_l, ranks_l = [], [], []
    for i in range(count):
        dims = [int(x) for x in rng.integers(16, 65, 4)]
        ranks = [1] + [int(x) for x in rng.integers(4, 33, 3)] + [1]
        tensors.append(eng.generate(dims, ranks, std=1.0, noise_rel=1e-5, seed=DATA_SEED + 1 + i))
        dims_l.append(dims)
        ranks_l.append(reachable_ranks(dims, ranks))   # a drawn rank above min(m, n) is unreachable
    elems = sum(int(np.prod(d)) for d in dims_l)
    kw = dict(s=10, q=2, precision="bf16", graph=True, autotune=True, streams=streams)
    out, info = eng.batch_tt_svd(tensors, dims_l, ranks_l, per_tensor=not grouped_only, **kw)   # tune, plan, capture
    bufs = info["core_buffers"]
    ts = []
    for _ in range(reps):
        out, info2 = eng.batch_tt_svd(tensors, dims_l, ranks_l, cores=bufs, **kw)
        ts.append(info2["total_ms"])
    t = float(np.median(ts))
    errs = [eng.rel_error(c, d, r, A) for c, d, r, A in zip(out[:32], dims_l[:32], ranks_l[:32], tensors[:32])]
    balg = batch_bytes_alg(dims_l, ranks_l)
    # the worker-stream engine on the same batch (the same tuned plan; the second call is timed)
    info_s = {"total_ms": float("nan"), "tensors_per_s": float("nan")}
    if not grouped_only:
        kws = dict(s=10, q=2, precision="bf16", streams=streams, graph=False, autotune=True, engine="streams")
        eng.batch_tt_svd(tensors, dims_l, ranks_l, **kws)
        _, info_s = eng.batch_tt_svd(tensors, dims_l, ranks_l, cores=bufs, **kws)
    res = {"workload": "512 order-4 TT tensors, N_i~U{16..64}, r_i~U{4..32}, N(0,1) cores + 1e-5 noise, BF16, s=10, q=2",
           "engine": info2["engine"], "launches_per_batch": info2["launches"],
           "tensors_per_s": round(count / (t * 1e-3), 1), "batch_ms": round(t, 3),
           "batch_ms_runs": [round(x, 3) for x in ts],
           "input_GBs": round(elems * 4 / (t * 1e-3) / 1e9, 1), "total_elements": elems,
           "bytes_alg": balg, "roofline": {"bound": "hbm", "achieved": round(balg / (t * 1e-3) / 1e9, 1),
                                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                           "frac": round(balg / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
           "p50_ms": round(info["p50_ms"], 4), "p99_ms": round(info["p99_ms"], 4),
           "graph": info2["graph"], "graph_reused": info2["graph_reused"], "sketch_chunks": info2["sketch_chunks"],
           "max_rel_error_first32": max(errs),
           "worker_streams": {"streams": streams, "batch_ms": round(info_s["total_ms"], 3),
                              "tensors_per_s": round(info_s["tensors_per_s"], 1)},
           "speedup_vs_worker_streams": round(info_s["total_ms"] / t, 2)}
    del tensors, out, bufs
    return res


def host_info():
    """CPU model and core counts of the host the CPU legs run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": usable, "cpu_count": os.cpu_count()}


def median_time(fn, runs=5):
    """one warm-up, then the median of runs timed calls"""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(A_dev, torch):
    """The reference algorithm of the order-6 drivers (exact Gram + eigensolver per step,
    ttdec6_singleGPU.cu:510-536) as the FP64 oracle (numpy, its BLAS threads) on a bounded sample of the
    same tensor: the first 4 slices along mode 6 (32^5 x 4); one warm-up, median of 5 runs."""
    from oracle import ttsvd
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        threads = 1
    n6 = 4
    sample = A_dev[: 32 ** 5 * n6].cpu().numpy()
    dims = DIMS[:5] + [n6]
    ranks = RANKS[:5] + [min(32, n6), 1]
    t, ts = median_time(lambda: ttsvd.tt_svd(sample, dims, ranks, method="exact"))
    gbs = sample.nbytes / t / 1e9
    out = {"value": round(gbs, 4), "unit": "GB/s", "cores": int(threads), "kind": "port",
           "sample": "oracle/ttsvd.py tt_svd(method='exact', FP64 numpy) on the first 4 mode-6 slices "
                     "(32^5 x 4 = %.1f MB) of the same tensor; median of 5 runs after a warm-up: %.3f s "
                     "(runs %s)" % (sample.nbytes / 1e6, t, ", ".join("%.3f" % x for x in ts))}
    out.update(host_info())
    return out


def cfg1_leg(eng, torch, L):
    """The reference's CPU-path configuration: plain third-order TT-SVD in FP32 of 256^3 (G1 1x256x16,
    G2 16x256x16, G3 16x256x1, N(0,1), + 1e-6 rms noise), ranks (16, 16), s = 8, q = 2, both steps
    randomized (EDrSVD, ttdec_gpu_float_rsvd_lu_half.cu:557-1014).  CPU: the C / OpenMP restatement
    (oracle/cpu_edrsvd.c: FP32 data, FP64 accumulation and solves), one warm-up and the median of 5 runs;
    GPU: the engine in FP32 (fixed q = 2) and in BF16 (the headline arithmetic, fixed q = 2)."""
    from oracle import cpu, ttsvd
    dims, ranks, s, q = [256] * 3, [1, 16, 16, 1], 8, 2
    A = eng.generate(dims, ranks, std=1.0, noise_rel=1e-6, seed=DATA_SEED)
    hA = A.cpu().numpy()
    oms = cpu.omegas(dims, ranks, s, seed=OMEGA_SEED)
    t_cpu, ts = median_time(lambda: cpu.tt_svd(hA, dims, ranks, s, q, oms))
    cores, _ = cpu.tt_svd(hA, dims, ranks, s, q, oms)
    shapes = ttsvd.tt_core_shapes(dims, ranks)
    e_cpu = ttsvd.rel_error(hA, ttsvd.tt_full([c.reshape(a, b, order="F") for c, (a, b) in zip(cores, shapes)],
                                              dims, ranks))
    gpu = {}
    for prec in ("fp32", "bf16"):
        kw = dict(s=s, q=q, seed=OMEGA_SEED, precision=prec)
        bufs = eng.core_buffers(dims, ranks, **kw)
        for _ in range(3):
            eng.tt_svd(A, dims, ranks, cores=bufs, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            c, _, rep = eng.tt_svd(A, dims, ranks, cores=bufs, **kw)
        torch.cuda.synchronize()
        tg = (time.perf_counter() - t0) / 20
        gpu[prec] = {"ms": round(tg * 1e3, 4), "GBs": round(hA.nbytes / tg / 1e9, 2),
                     "rel_error": eng.rel_error(c, dims, ranks, A), "paths": rep["path"]}
    out = {"workload": "third-order TT-SVD 256^3 (67.1 MB FP32), ranks (16,16), s=8, q=2, Gaussian sketch",
           "cpu": {"ms": round(t_cpu * 1e3, 3), "GBs": round(hA.nbytes / t_cpu / 1e9, 4), "rel_error": e_cpu,
                   "threads": cpu.threads(), "kind": "port (C / OpenMP restatement, oracle/cpu_edrsvd.c)",
                   "runs_ms": [round(x * 1e3, 3) for x in ts]},
           "gpu_fp32": gpu["fp32"], "gpu_bf16": gpu["bf16"],
           "speedup_gpu_fp32_vs_cpu": round(t_cpu / (gpu["fp32"]["ms"] * 1e-3), 1)}
    out["cpu"].update(host_info())
    del A
    return out


def copy_probe(torch):
    """one 1 GiB device-to-device copy: the scale reference of rocprof's FETCH_SIZE / WRITE_SIZE"""
    a = torch.ones(1 << 28, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported (the multi-GPU "
              "sharding is a design, not a build)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    from tensor_train_ring_decompositions_hpc_gpu_amd.api import Engine

    torch.cuda.set_device(0)
    eng = Engine()
    side = {"tr": lambda: tr_leg(eng, torch, L), "batch": lambda: batch_leg(eng, torch, grouped_only=args.batch_grouped_only),
            "learned": lambda: learned_leg(eng, torch, L), "cfg1": lambda: cfg1_leg(eng, torch, L)}
    if args.only in side:
        # one secondary workload alone (profiling runs): its own JSON object, not the bench line
        res = side[args.only]()
        if args.copy_probe:
            copy_probe(torch)
        print(json.dumps({args.only: res}))
        eng.close()
        return
    A = eng.generate(DIMS, RANKS, std=1.0 / np.sqrt(32.0), noise_rel=1e-6, seed=DATA_SEED)
    torch.cuda.synchronize()
    in_bytes = A.numel() * 4
    flags = L.FLAG_FAST_RANGE if args.fast_range else 0
    if not args.tail_x4:
        flags |= L.FLAG_TAIL_X3
    if not args.no_adaptive:
        flags |= L.FLAG_ADAPTIVE

    legs = {}
    if args.only in ("", "bf16"):
        legs["bf16"] = run_leg(eng, A, torch, "bf16", args.q, args.s, args.steps, args.warmup, flags,
                               rank_mode=args.rank_mode)
    if args.only == "" and args.rank_mode == "learned":
        legs["bf16_fixed"] = run_leg(eng, A, torch, "bf16", args.q, args.s, max(1, args.steps // 2), 1, flags)
    if args.only == "":
        # the schedule as written (fixed q = 2, no a-posteriori check) in BF16, and the adaptive schedule
        # in FP32: with the two legs below, the speedup splits into arithmetic and schedule
        legs["bf16_q2"] = run_leg(eng, A, torch, "bf16", 2, args.s, max(1, args.steps // 2), 1,
                                  flags & ~L.FLAG_ADAPTIVE)
        legs["fp32_adaptive"] = run_leg(eng, A, torch, "fp32", args.q, args.s, max(1, args.steps // 2), 1,
                                        L.FLAG_ADAPTIVE)
    if args.only in ("", "fp32"):
        legs["fp32"] = run_leg(eng, A, torch, "fp32", 2, 10, max(1, args.steps // 2), 1)
    head = legs.get("bf16") or legs["fp32"]
    t, err, rep, sig = head
    gbs = in_bytes / t / 1e9

    # dominant kernel class of the headline run and its roofline
    ks = rep["kernels"]
    # dominant streaming (HBM-bound) kernel class; the FP64 small-matrix kernels move no HBM bytes.
    # The roofline is quoted on that class's largest launch (its algorithmic bytes over its device
    # time from hipEvents on the engine's stream); traffic from the committed rocprof counters
    stream = {k: v for k, v in ks.items() if v["bytes"] > 0}
    dom = max(stream, key=lambda k: stream[k]["ms"])
    kd = ks[dom]
    achieved = kd["top_bytes"] / (kd["top_ms"] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "tt6_32_r32_bf16_adaptive_traffic.json")
    if os.path.exists(tf) and "bf16" in legs and not args.no_adaptive:
        tk = json.load(open(tf))["kernels"].get(dom)
        if tk:
            traffic = tk["traffic_bytes"]
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": kd["top_bytes"], "ms_per_launch": round(kd["top_ms"], 4),
            "class_total": {"launches": kd["launches"], "ms": round(kd["ms"], 4),
                            "GBs": round(kd["bytes"] / (kd["ms"] * 1e-3) / 1e9, 1)}}
    pipeline_frac = (rep["bytes_alg"] / (HBM_PEAK_GBS * 1e9)) / t

    out = {
        "metric": METRIC, "value": round(gbs, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16" if "bf16" in legs else "fp32",
        "data": "synthetic: seeded TT cores N(0,1/32) ranks (1,32,32,32,32,32,1) contracted + 1e-6 relative noise",
        "config": {"workload": "order-6 TT-SVD 32^6 (4.295 GB FP32), %s, s=%d, q%s%d" % (
                       "learned ranks (ceiling 32, eps=1e-4)" if args.rank_mode == "learned" and "bf16" in legs
                       else "ranks 32", args.s, "=" if args.no_adaptive else "<=", args.q),
                   "dims": DIMS, "ranks": RANKS, "oversampling": args.s, "power_iters": args.q,
                   "rank_mode": args.rank_mode if "bf16" in legs else "fixed",
                   "eps": 1e-4, "ranks_used": rep["ranks"], "predicted_rank": rep["predicted_rank"],
                   "range_arith": "bf16" if args.fast_range else "bf16x3",
                   "projection_arith": "bf16x4" if args.tail_x4 else "bf16x3 (randomized steps' Z = A^T Q)",
                   "range_check": None if args.no_adaptive else "a-posteriori, 8 test columns, eps=1e-4",
                   "paths": rep["path"], "ell": rep["ell"], "power_iters_run": rep["power_iters"],
                   "passes": rep["passes"], "range_resid": rep["range_resid"]},
        "rel_error": err,
        "pipeline_hbm_frac": round(pipeline_frac, 4),
        "bytes_alg": rep["bytes_alg"],
        "kernels": {k: {"ms": round(v["ms"], 4), "launches": v["launches"],
                        "GBs": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in ks.items()},
        "roofline": roof,
    }
    if "bf16_fixed" in legs:
        tf_, ef_, rf_, _ = legs["bf16_fixed"]
        out["bf16_fixed_ranks"] = {"value": round(in_bytes / tf_ / 1e9, 2), "ms_per_step": round(tf_ * 1e3, 4),
                                   "rel_error": ef_, "passes": rf_["passes"]}
    for name, key, desc in (("bf16_q2", "bf16_fixed_q2", "BF16 (bf16x3 range finder, bf16x3 projection), fixed "
                                                           "ranks 32, q = 2 fixed, no range check: cfg2's schedule as written"),
                            ("fp32_adaptive", "fp32_adaptive", "FP32 MFMA, fixed ranks 32, a-posteriori check, q <= 2")):
        if name in legs:
            tl, el, rl, _ = legs[name]
            out[key] = {"value": round(in_bytes / tl / 1e9, 2), "ms_per_step": round(tl * 1e3, 4), "rel_error": el,
                        "passes": rl["passes"], "power_iters_run": rl["power_iters"], "bytes_alg": rl["bytes_alg"],
                        "pipeline_hbm_frac": round(rl["bytes_alg"] / (HBM_PEAK_GBS * 1e9) / tl, 4), "leg": desc}
    if "fp32" in legs and "bf16" in legs:
        t32, err32, rep32, _ = legs["fp32"]
        out["fp32_plain"] = {"value": round(in_bytes / t32 / 1e9, 2), "ms_per_step": round(t32 * 1e3, 4),
                             "rel_error": err32, "power_iters": 2, "oversampling": 10,
                             "passes": rep32["passes"], "bytes_alg": rep32["bytes_alg"]}
        out["speedup_vs_fp32_plain"] = round(t32 / t, 3)
        if "bf16_q2" in legs and "fp32_adaptive" in legs:
            # arithmetic: FP32 -> BF16 on the same schedule; schedule: fixed q = 2 -> adaptive
            out["speedup_split"] = {
                "arithmetic_at_q2": round(t32 / legs["bf16_q2"][0], 3),
                "arithmetic_adaptive": round(legs["fp32_adaptive"][0] / legs["bf16_fixed"][0], 3)
                if "bf16_fixed" in legs else None,
                "schedule_fp32": round(t32 / legs["fp32_adaptive"][0], 3),
                "schedule_bf16": round(legs["bf16_q2"][0] / legs["bf16_fixed"][0], 3) if "bf16_fixed" in legs else None,
                "learned_vs_fixed_ranks": round(legs["bf16_fixed"][0] / t, 3) if "bf16_fixed" in legs else None}
    if not args.no_learned and args.only == "":
        out["learned_rank"] = learned_leg(eng, torch, L)
    if not args.no_tr and args.only == "":
        out["tr_fp16"] = tr_leg(eng, torch, L)
    if not args.no_batch and args.only == "":
        out["batched"] = batch_leg(eng, torch)
    if not args.no_host and args.only == "":
        out["host_streamed"] = host_leg(eng, A, torch, L)
    if not args.no_cpu and args.only == "":
        out["cfg1_third_order"] = cfg1_leg(eng, torch, L)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(A, torch)
    if args.copy_probe:
        copy_probe(torch)
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
