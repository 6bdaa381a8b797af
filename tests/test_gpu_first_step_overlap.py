"""The first step's eigensolve beside the second step.

A step that follows a deferred full-rank step (square orthogonal core P) runs on the input S itself:
A = (I (x) P^T) S and S give the same sketch Gram, range residual, Z, eigenvectors and remainder, and
P enters only the core, Q_A = (I (x) P^T) Q_S.  When a BF16 TT sweep's first pass takes both the
exact 32-row first step's Gram and the second step's sketch, the first step's eigensolve and core
then run on a side stream beside the second step, which joins before its rank decision.
TTX_FLAG_TIMING keeps everything on one stream, so the two orders run the same kernels on the same
inputs and must agree bit for bit:

* full first-step rank (fixed and DELTA ranks);
* first-step DELTA rank below 32: the speculative second step is discarded and the sweep projects the
  first step and runs the second on its remainder, as the single-stream order does.
"""
import numpy as np
import pytest
import torch

from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L

pytestmark = pytest.mark.gpu

DIMS = [32, 32, 32, 32]


def _run(engine, A, ranks, timing, **kw):
    flags = L.FLAG_ADAPTIVE | L.FLAG_TAIL_X3 | (L.FLAG_TIMING if timing else 0)
    cores, sig, rep = engine.tt_svd(A, DIMS, ranks, s=10, q=2, precision="bf16", flags=flags, **kw)
    torch.cuda.synchronize()
    err = engine.rel_error(cores, DIMS, rep["ranks"], A)
    host = [c[: rows * cols].detach().cpu().numpy().copy() for c, (rows, cols) in cores]
    return host, [np.asarray(s).copy() for s in sig], rep, err


@pytest.mark.parametrize("mode", ["fixed", "delta"])
def test_overlapped_first_step_matches_serial(engine, mode):
    true = [1, 32, 32, 32, 1]
    A = engine.generate(DIMS, true, std=32 ** -0.5, noise_rel=1e-6, seed=21)
    kw = dict(rank_mode=mode, eps=1e-4) if mode == "delta" else {}
    print("overlap vs serial:", mode, flush=True)
    c1, s1, r1, e1 = _run(engine, A, true, False, **kw)
    c2, s2, r2, e2 = _run(engine, A, true, True, **kw)
    assert r1["ranks"] == r2["ranks"] == true, (r1["ranks"], r2["ranks"])
    assert r1["path"] == r2["path"] and r1["passes"] == r2["passes"]
    assert abs(e1 - e2) <= 1e-9 * e2 and e1 < 2e-5, (e1, e2)
    for k in range(4):
        np.testing.assert_array_equal(c1[k], c2[k])
    for k in range(3):
        np.testing.assert_array_equal(s1[k], s2[k])


def test_first_step_rank_below_ceiling_reruns_second_step(engine):
    true = [1, 8, 8, 8, 1]
    A = engine.generate(DIMS, true, std=1.0, noise_rel=1e-6, seed=22)
    ceil = [1, 32, 32, 32, 1]
    print("overlap retry: delta ranks below the ceiling", flush=True)
    c1, s1, r1, e1 = _run(engine, A, ceil, False, rank_mode="delta", eps=1e-3)
    c2, s2, r2, e2 = _run(engine, A, ceil, True, rank_mode="delta", eps=1e-3)
    assert r1["ranks"][1] == 8, r1["ranks"]
    assert r1["ranks"] == r2["ranks"]
    assert abs(e1 - e2) <= 1e-9 * e2 and e1 < 1e-3, (e1, e2)
    for k in range(4):
        np.testing.assert_array_equal(c1[k], c2[k])
    for k in range(3):
        np.testing.assert_array_equal(s1[k], s2[k])
