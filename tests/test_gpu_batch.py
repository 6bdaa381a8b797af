"""Batched TT-SVD (ttx_batch_tt_svd) on both engines.  Grouped engine (one launch per kernel variant
and step phase for the whole batch): every tensor's cores equal, bit for bit, those of ttx_tt_svd on
that tensor alone with TTX_FLAG_BATCH_PLAN (the same kernels and arguments, the blocks of all the
tensors' launches concatenated), eager and captured, on every replay; the worker-stream engine gives
the same cores again; the cores reconstruct each tensor as the FP64 oracle's do."""
import numpy as np
import pytest

from oracle import ttsvd

pytestmark = pytest.mark.gpu


def _batch(engine, count, seed, dims_range=(8, 24), rank_range=(2, 8)):
    """count order-4 tensors with drawn sizes and ranks; a drawn rank above min(m, n) cannot be reached
    (the engine keeps min(m, n) columns and leaves the rest of a caller's buffer unwritten), so the
    batch asks for the reachable ranks"""
    from tensor_train_ring_decompositions_hpc_gpu_amd.api import reachable_ranks
    rng = np.random.default_rng(seed)
    tensors, dims_l, ranks_l = [], [], []
    for i in range(count):
        dims = [int(x) for x in rng.integers(dims_range[0], dims_range[1] + 1, 4)]
        ranks = [1] + [int(x) for x in rng.integers(rank_range[0], rank_range[1] + 1, 3)] + [1]
        tensors.append(engine.generate(dims, ranks, std=1.0, noise_rel=1e-5, seed=1000 + i))
        dims_l.append(dims)
        ranks_l.append(reachable_ranks(dims, ranks))
    return tensors, dims_l, ranks_l


def _host(cores):
    return [c.detach().cpu().numpy().copy() for c, _ in cores]


def _single(engine, A, dims, ranks, **kw):
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    flags = kw.pop("flags", 0) | L.FLAG_BATCH_PLAN
    ref, _, rep = engine.tt_svd(A, dims, ranks, flags=flags, **kw)
    return ref, rep


@pytest.mark.parametrize("eng,graph", [("grouped", False), ("grouped", True), ("streams", False), ("streams", True)])
def test_batch_matches_single_calls(engine, eng, graph):
    tensors, dims_l, ranks_l = _batch(engine, 12, 3)
    print("batch engine=%s graph=%s" % (eng, graph), flush=True)
    out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=6, q=1, precision="bf16", streams=3, graph=graph,
                                    autotune=False, engine=eng)
    assert info["graph"] == int(graph) and info["engine"] == eng
    for A, dims, ranks, cores in zip(tensors, dims_l, ranks_l, out):
        ref, _ = _single(engine, A, dims, ranks, s=6, q=1, precision="bf16")
        for (c, _), (r, _) in zip(cores, ref):
            assert np.array_equal(c.cpu().numpy(), r.cpu().numpy())


def test_grouped_paths_and_shapes_match_single_calls(engine):
    """Every step path in one batch: exact Gram steps (m <= 64 with r + s >= m, including a full-rank
    one), exact short-side steps (n <= 64, m > 64), randomized steps with 32- and 64-column sketches and
    q = 2 power iterations; FP32 and BF16 (three-product tail); each tensor bit-exact against its own
    batch-plan sweep, and within the BF16 tolerance of the FP64 oracle's reconstruction error."""
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    from tensor_train_ring_decompositions_hpc_gpu_amd.api import reachable_ranks
    shapes = [([16, 20, 24, 12], [1, 16, 20, 8, 1]),      # r1 = N1 = 16: full-rank exact first step
              ([48, 40, 16, 70], [1, 12, 30, 24, 1]),     # randomized first step (ell 22 < 48)
              ([64, 32, 40, 20], [1, 28, 32, 12, 1]),     # 64-column sketch (ell 38), exact-n last step
              ([24, 64, 64, 8], [1, 8, 6, 4, 1]),
              ([33, 17, 29, 45], [1, 5, 31, 9, 1]),
              ([40, 40, 40, 40], [1, 32, 32, 32, 1])]
    tensors, dims_l, ranks_l = [], [], []
    for i, (dims, ranks) in enumerate(shapes):
        tensors.append(engine.generate(dims, ranks, std=1.0, noise_rel=1e-5, seed=2000 + i))
        dims_l.append(dims)
        ranks_l.append(reachable_ranks(dims, ranks))
    for prec, flags in (("fp32", 0), ("bf16", L.FLAG_TAIL_X3)):
        print("grouped paths precision=%s" % prec, flush=True)
        out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=10, q=2, precision=prec, flags=flags,
                                        graph=False, autotune=False)
        assert info["engine"] == "grouped" and info["launches"] > 0
        paths = set()
        for A, dims, ranks, cores in zip(tensors, dims_l, ranks_l, out):
            ref, rep = _single(engine, A, dims, ranks, s=10, q=2, precision=prec, flags=flags)
            paths.update(rep["path"])
            for (c, _), (r, _) in zip(cores, ref):
                assert np.array_equal(c.cpu().numpy(), r.cpu().numpy()), (prec, dims)
            err = engine.rel_error(cores, dims, ranks, A)
            An = A.cpu().numpy().astype(np.float64)
            c_ref, _, _, used = ttsvd.tt_svd(An, dims, ranks, s=10, q=2)
            err_ref = ttsvd.rel_error(An, ttsvd.tt_full(c_ref, dims, used))
            assert abs(err - err_ref) <= 1e-5 + 2e-2 * err_ref, (prec, dims, err, err_ref)
        assert paths == {0, 1, 2}, paths


def test_graph_replays_match_eager(engine):
    import torch
    tensors, dims_l, ranks_l = _batch(engine, 40, 5, dims_range=(8, 32), rank_range=(2, 12))
    eager, _ = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=6, q=1, graph=False, autotune=False)
    ref = [_host(c) for c in eager]
    out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=6, q=1, graph=True, autotune=False)
    bufs = info["core_buffers"]
    assert info["graph_reused"] == 0
    for rep in range(3):
        # eager work and synchronisation between the replays
        junk = torch.randn(1 << 20, device="cuda")
        engine.tt_svd(tensors[0], dims_l[0], ranks_l[0], s=6, q=1)
        for cs in bufs:
            for c in cs:
                c.fill_(float("nan"))
        torch.cuda.synchronize()
        out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=6, q=1, graph=True, autotune=False, cores=bufs)
        assert info["graph_reused"] == 1
        for r, cores in zip(ref, out):
            for a, b in zip(r, _host(cores)):
                assert np.array_equal(a, b), rep
        del junk


def test_grouped_plan_rebuilt_for_another_batch(engine):
    """A different batch in between replaces the plan (and the arena the first graph referred to): the
    first batch then builds and captures again (graph_reused == 0) and still gives its eager result."""
    small, dims_s, ranks_s = _batch(engine, 6, 11, dims_range=(8, 12), rank_range=(2, 4))
    ref_out, _ = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, graph=False, autotune=False)
    ref = [_host(c) for c in ref_out]
    out, info = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, graph=True, autotune=False)
    bufs = info["core_buffers"]
    big, dims_b, ranks_b = _batch(engine, 4, 12, dims_range=(24, 32), rank_range=(8, 12))
    engine.batch_tt_svd(big, dims_b, ranks_b, s=10, q=1, graph=True, autotune=False)
    out, info = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, graph=True, autotune=False, cores=bufs)
    assert info["graph"] == 1 and info["graph_reused"] == 0
    for r, cores in zip(ref, out):
        for a, b in zip(r, _host(cores)):
            assert np.array_equal(a, b)


def test_batch_reconstructs_like_oracle(engine):
    tensors, dims_l, ranks_l = _batch(engine, 16, 9, dims_range=(16, 40), rank_range=(4, 12))
    for eng in ("grouped", "streams"):
        out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=10, q=2, streams=4, graph=True, autotune=True,
                                        per_tensor=True, engine=eng)
        assert info["p50_ms"] > 0 and info["p99_ms"] >= info["p50_ms"]
        assert info["tuned_shapes"] >= 1 and info["sketch_chunks"] in (8, 16, 32, 64)
        for A, dims, ranks, cores in zip(tensors, dims_l, ranks_l, out):
            err = engine.rel_error(cores, dims, ranks, A)
            An = A.cpu().numpy().astype(np.float64)
            c_ref, _, _, used = ttsvd.tt_svd(An, dims, ranks, s=10, q=2)
            err_ref = ttsvd.rel_error(An, ttsvd.tt_full(c_ref, dims, used))
            assert abs(err - err_ref) <= 1e-5 + 2e-2 * err_ref, (eng, err, err_ref)


def test_worker_stream_graph_refuses_batches_above_its_tested_size(engine):
    """TTX_BATCH_GRAPH_MAX: a worker-stream graph batch above the size its tests cover is refused before
    any work is queued; the grouped engine captures the same batch."""
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    tensors, dims_l, ranks_l = _batch(engine, 17, 21, dims_range=(8, 16), rank_range=(2, 4))
    with pytest.raises(L.TtxError, match="at most 16"):
        engine.batch_tt_svd(tensors, dims_l, ranks_l, s=4, q=1, streams=2, graph=True, autotune=False,
                            engine="streams")
    out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=4, q=1, graph=True, autotune=False)
    assert info["graph"] == 1 and info["engine"] == "grouped" and len(out) == 17


def test_worker_graph_not_relaunched_after_worker_buffers_moved(engine):
    """A larger eager batch on the same handle reallocates the workers' buffers; the captured graph of
    the first batch refers to the old ones, so the next identical graph batch must capture again
    (graph_reused == 0) and still give the eager result bit for bit.  A fresh handle, so that the
    workers' buffers start at the small batch's sizes."""
    import torch
    from tensor_train_ring_decompositions_hpc_gpu_amd.api import Engine
    fresh = Engine()
    try:
        _graph_after_growth(fresh)
    finally:
        torch.cuda.synchronize()
        fresh.close()


def _graph_after_growth(engine):
    kw = dict(streams=2, autotune=False, engine="streams")
    small, dims_s, ranks_s = _batch(engine, 6, 11, dims_range=(8, 12), rank_range=(2, 4))
    ref_out, _ = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, graph=False, **kw)
    ref = [_host(c) for c in ref_out]
    out, info = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, graph=True, **kw)
    bufs = info["core_buffers"]
    big, dims_b, ranks_b = _batch(engine, 4, 12, dims_range=(24, 32), rank_range=(8, 12))
    engine.batch_tt_svd(big, dims_b, ranks_b, s=10, q=1, graph=False, **kw)
    out, info = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, graph=True, cores=bufs, **kw)
    assert info["graph"] == 1 and info["graph_reused"] == 0
    for r, cores in zip(ref, out):
        for a, b in zip(r, _host(cores)):
            assert np.array_equal(a, b)
    # with nothing reallocated since, the new capture is relaunched
    out, info = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, graph=True, cores=bufs, **kw)
    assert info["graph_reused"] == 1
    for r, cores in zip(ref, out):
        for a, b in zip(r, _host(cores)):
            assert np.array_equal(a, b)


def test_fp16_batch_runs_on_worker_streams(engine):
    tensors, dims_l, ranks_l = _batch(engine, 4, 17)
    out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=6, q=1, precision="fp16", graph=False,
                                    autotune=False)
    assert info["engine"] == "streams"
    for A, dims, ranks, cores in zip(tensors, dims_l, ranks_l, out):
        ref, _ = _single(engine, A, dims, ranks, s=6, q=1, precision="fp16")
        for (c, _), (r, _) in zip(cores, ref):
            assert np.array_equal(c.cpu().numpy(), r.cpu().numpy())


def test_batch_refuses_adaptive_range_finder(engine):
    """The range check reads its residual back to the host between passes: batches (eager or captured,
    either engine) refuse TTX_FLAG_ADAPTIVE before queueing any work."""
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    tensors, dims_l, ranks_l = _batch(engine, 3, 13)
    for eng in ("grouped", "streams"):
        for graph in (False, True):
            with pytest.raises(L.TtxError, match="adaptive"):
                engine.batch_tt_svd(tensors, dims_l, ranks_l, s=4, q=1, streams=2, graph=graph, autotune=False,
                                    flags=L.FLAG_ADAPTIVE, engine=eng)


def test_grouped_reports_eigensolver_cap(monkeypatch):
    """A one-sweep cap on the Jacobi eigensolver (TTX_JACOBI_MAX_SWEEPS, read at handle creation) is
    reported by the grouped engine as TTX_ERR_NO_CONVERGE after the batch ran."""
    import torch
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    from tensor_train_ring_decompositions_hpc_gpu_amd.api import Engine
    monkeypatch.setenv("TTX_JACOBI_MAX_SWEEPS", "1")
    eng = Engine()
    monkeypatch.delenv("TTX_JACOBI_MAX_SWEEPS")
    try:
        tensors, dims_l, ranks_l = _batch(eng, 3, 19, dims_range=(16, 24), rank_range=(6, 10))
        with pytest.raises(L.TtxError) as ei:
            eng.batch_tt_svd(tensors, dims_l, ranks_l, s=6, q=1, graph=False, autotune=False)
        assert ei.value.status == L.ERR_NO_CONVERGE
    finally:
        torch.cuda.synchronize()
        eng.close()
