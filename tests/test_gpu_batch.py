"""Batched TT-SVD (ttx_batch_tt_svd): the batch, its captured graph and every replay of that graph give
the same cores as one ttx_tt_svd call per tensor; the cores reconstruct each tensor as the FP64 oracle's
do."""
import numpy as np
import pytest

from oracle import ttsvd

pytestmark = pytest.mark.gpu


def _batch(engine, count, seed, dims_range=(8, 24), rank_range=(2, 8)):
    rng = np.random.default_rng(seed)
    tensors, dims_l, ranks_l = [], [], []
    for i in range(count):
        dims = [int(x) for x in rng.integers(dims_range[0], dims_range[1] + 1, 4)]
        ranks = [1] + [int(x) for x in rng.integers(rank_range[0], rank_range[1] + 1, 3)] + [1]
        tensors.append(engine.generate(dims, ranks, std=1.0, noise_rel=1e-5, seed=1000 + i))
        dims_l.append(dims)
        ranks_l.append(ranks)
    return tensors, dims_l, ranks_l


def _host(cores):
    return [c.detach().cpu().numpy().copy() for c, _ in cores]


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_calls(engine, graph):
    tensors, dims_l, ranks_l = _batch(engine, 12, 3)
    out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=6, q=1, precision="bf16", streams=3, graph=graph,
                                    autotune=False)
    assert info["graph"] == int(graph)
    for A, dims, ranks, cores in zip(tensors, dims_l, ranks_l, out):
        ref, _, _ = engine.tt_svd(A, dims, ranks, s=6, q=1, precision="bf16")
        for (c, _), (r, _) in zip(cores, ref):
            # same kernels and plans in the same order: identical results
            assert np.array_equal(c.cpu().numpy(), r.cpu().numpy())


def test_graph_replays_match_eager(engine):
    import torch
    tensors, dims_l, ranks_l = _batch(engine, 10, 5)
    eager, _ = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=6, q=1, streams=2, graph=False, autotune=False)
    ref = [_host(c) for c in eager]
    out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=6, q=1, streams=2, graph=True, autotune=False)
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
        out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=6, q=1, streams=2, graph=True, autotune=False,
                                        cores=bufs)
        assert info["graph_reused"] == 1
        for r, cores in zip(ref, out):
            for a, b in zip(r, _host(cores)):
                assert np.array_equal(a, b), rep
        del junk


def test_batch_autotuned_reconstructs(engine):
    tensors, dims_l, ranks_l = _batch(engine, 16, 9, dims_range=(16, 40), rank_range=(4, 12))
    out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=10, q=2, streams=4, graph=True, autotune=True,
                                    per_tensor=True)
    assert info["tuned_shapes"] >= 1 and info["p50_ms"] > 0 and info["p99_ms"] >= info["p50_ms"]
    for A, dims, ranks, cores in zip(tensors, dims_l, ranks_l, out):
        err = engine.rel_error(cores, dims, ranks, A)
        An = A.cpu().numpy().astype(np.float64)
        c_ref, _, _, used = ttsvd.tt_svd(An, dims, ranks, s=10, q=2)
        err_ref = ttsvd.rel_error(An, ttsvd.tt_full(c_ref, dims, used))
        assert abs(err - err_ref) <= 1e-4 + 2e-2 * err_ref, (err, err_ref)


def test_graph_capture_refuses_batches_above_its_tested_size(engine):
    """TTX_BATCH_GRAPH_MAX: a graph batch above the size the graph tests cover is refused before any
    work is queued; the same batch runs without the graph."""
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    tensors, dims_l, ranks_l = _batch(engine, 17, 21, dims_range=(8, 16), rank_range=(2, 4))
    with pytest.raises(L.TtxError, match="at most 16"):
        engine.batch_tt_svd(tensors, dims_l, ranks_l, s=4, q=1, streams=2, graph=True, autotune=False)
    out, info = engine.batch_tt_svd(tensors, dims_l, ranks_l, s=4, q=1, streams=2, graph=False, autotune=False)
    assert info["graph"] == 0 and len(out) == 17


def test_graph_not_relaunched_after_worker_buffers_moved(engine):
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
    small, dims_s, ranks_s = _batch(engine, 6, 11, dims_range=(8, 12), rank_range=(2, 4))
    ref_out, _ = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, streams=2, graph=False, autotune=False)
    ref = [_host(c) for c in ref_out]
    out, info = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, streams=2, graph=True, autotune=False)
    bufs = info["core_buffers"]
    big, dims_b, ranks_b = _batch(engine, 4, 12, dims_range=(24, 32), rank_range=(8, 12))
    engine.batch_tt_svd(big, dims_b, ranks_b, s=10, q=1, streams=2, graph=False, autotune=False)
    out, info = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, streams=2, graph=True, autotune=False,
                                    cores=bufs)
    assert info["graph"] == 1 and info["graph_reused"] == 0
    for r, cores in zip(ref, out):
        for a, b in zip(r, _host(cores)):
            assert np.array_equal(a, b)
    # with nothing reallocated since, the new capture is relaunched
    out, info = engine.batch_tt_svd(small, dims_s, ranks_s, s=4, q=1, streams=2, graph=True, autotune=False,
                                    cores=bufs)
    assert info["graph_reused"] == 1
    for r, cores in zip(ref, out):
        for a, b in zip(r, _host(cores)):
            assert np.array_equal(a, b)


def test_batch_refuses_adaptive_range_finder(engine):
    """The range check reads its residual back to the host between passes: batches (eager or captured)
    refuse TTX_FLAG_ADAPTIVE before queueing any work."""
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    tensors, dims_l, ranks_l = _batch(engine, 3, 13)
    for graph in (False, True):
        with pytest.raises(L.TtxError, match="adaptive"):
            engine.batch_tt_svd(tensors, dims_l, ranks_l, s=4, q=1, streams=2, graph=graph, autotune=False,
                                flags=L.FLAG_ADAPTIVE)
