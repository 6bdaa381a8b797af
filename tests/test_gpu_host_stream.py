"""Host-streamed TT sweep (ttx_tt_svd_host), the counterpart of ttdec_half_lu_host_6
(ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:656-796, its slice loops :540-652): the tensor
stays in host memory and the first step's Gram and projection read it in column slices.

Pinned by: the first step's singular values against numpy's SVD of the unfolding (FP64), the whole
result against the device-resident sweep on the same tensor, exact-rank reconstruction through the
oracle's contraction, and the delta rule recovering the generating ranks."""
import numpy as np
import pytest
import torch

from oracle import ttsvd
from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L

pytestmark = pytest.mark.gpu

DIMS = [16, 12, 10, 8, 6]
RANKS = [1, 6, 8, 7, 5, 1]


def _cores_host(cores):
    return [c.detach().cpu().numpy() for c, _ in cores]


@pytest.mark.parametrize("twice", [False, True])   # one read into a device copy / two slice passes
@pytest.mark.parametrize("slice_cols", [0, 5760, 1000, 7])   # auto (one slice), exactly one, ragged, many
def test_host_stream_matches_device_sweep(engine, slice_cols, twice):
    A = engine.generate(DIMS, RANKS, std=1.0, noise_rel=1e-6, seed=7)
    Ah = A.cpu().pin_memory()
    print(f"host stream: dims {DIMS} ranks {RANKS} slice_cols {slice_cols} twice {twice}", flush=True)
    cd, sd, rd = engine.tt_svd(A, DIMS, RANKS, precision="fp32")
    flags = L.FLAG_STREAM_TWICE if twice else 0
    ch, sh, rh = engine.tt_svd_host(Ah, DIMS, RANKS, slice_cols=slice_cols, precision="fp32", flags=flags)
    torch.cuda.synchronize()
    assert rh["ranks"] == rd["ranks"] == RANKS
    assert rh["path"][0] == L.PATH_EXACT_M and rh["passes"][0] == (2 if twice else 1)
    # step 0 against numpy: the leading singular values of the 16 x 5760 unfolding (the noise
    # floor's values come from a Gram and carry its rounding, so they are not compared)
    a = A.cpu().numpy().astype(np.float64)
    A1 = a.reshape(-1, DIMS[0]).T
    s_np = np.linalg.svd(A1, compute_uv=False)
    r1 = RANKS[1]
    np.testing.assert_allclose(sh[0][:r1], s_np[:r1], rtol=0, atol=2e-6 * s_np[0])
    # every step's singular values agree with the device-resident sweep (exact ranks: the sketch
    # captures the whole range, so the values do not depend on the sketch)
    for k in range(len(DIMS) - 1):
        np.testing.assert_allclose(sh[k][:RANKS[k + 1]], sd[k][:RANKS[k + 1]], rtol=0, atol=2e-5 * sd[k][0])
    # the first core is the same eigenvector basis under the same sign gauge
    np.testing.assert_allclose(_cores_host(ch)[0], _cores_host(cd)[0], rtol=0, atol=2e-5)
    full = ttsvd.tt_full(_cores_host(ch), DIMS, rh["ranks"])
    e_h = ttsvd.rel_error(a, full)
    e_d = ttsvd.rel_error(a, ttsvd.tt_full(_cores_host(cd), DIMS, rd["ranks"]))
    assert e_h < 1e-5 and abs(e_h - e_d) < 2e-6
    assert abs(engine.rel_error(ch, DIMS, rh["ranks"], A) - e_h) < 1e-6


def test_host_stream_pageable_numpy_and_bf16(engine):
    """A pageable numpy array works too (synchronous staging), in the default BF16 arithmetic."""
    A = engine.generate(DIMS, RANKS, std=1.0, noise_rel=1e-6, seed=11)
    a = np.ascontiguousarray(A.cpu().numpy())
    ch, sh, rh = engine.tt_svd_host(a, DIMS, RANKS, slice_cols=999, flags=L.FLAG_STREAM_TWICE)
    torch.cuda.synchronize()
    assert rh["ranks"] == RANKS
    e = ttsvd.rel_error(a.astype(np.float64), ttsvd.tt_full(_cores_host(ch), DIMS, rh["ranks"]))
    assert e < 1e-4


def test_host_stream_delta_ranks(engine):
    """DELTA: the first rank from the exact spectrum, the rest with the same per-step budget; the
    generating ranks come back and the error stays within eps."""
    dims = [16, 12, 10, 8]
    true = [1, 5, 6, 4, 1]
    A = engine.generate(dims, true, std=1.0, noise_rel=1e-6, seed=5)
    Ah = A.cpu().pin_memory()
    ceil = [1, 16, 32, 8, 1]
    ch, sh, rh = engine.tt_svd_host(Ah, dims, ceil, slice_cols=100, rank_mode="delta", eps=1e-3,
                                    flags=L.FLAG_STREAM_TWICE)
    cd, sd, rd = engine.tt_svd(A, dims, ceil, rank_mode="delta", eps=1e-3)
    torch.cuda.synchronize()
    assert rh["ranks"] == true == rd["ranks"]
    e = ttsvd.rel_error(A.cpu().numpy().astype(np.float64), ttsvd.tt_full(_cores_host(ch), dims, rh["ranks"]))
    assert e < 1e-3


@pytest.mark.parametrize("twice", [False, True])
@pytest.mark.parametrize("dims,ranks", [
    ([8, 8, 8, 8, 8, 8], [2, 2, 8, 8, 8, 8, 2]),   # ring of ranks 2, inner ranks r0^2 r_k: exact
    ([4, 6, 6, 6], [2, 4, 8, 8, 2]),               # r0 r1 = 8 > N_1 = 4: the first core zero-padded
])
def test_tr_host_stream_matches_device_sweep(engine, dims, ranks, twice):
    """TR form (trdec6_singleGPU.cu:668-826): the streamed first step of rank r0 r1, the 231 relayout,
    then the sweep over r1 N2, .., Nd r0; compared with the device TR sweep and the oracle."""
    A = engine.generate(dims, [2] * (len(dims) + 1), std=1.0, noise_rel=1e-6, seed=1134, tr=True)
    Ah = A.cpu().pin_memory()
    print(f"tr host stream: dims {dims} ranks {ranks} twice {twice}", flush=True)
    flags = L.FLAG_STREAM_TWICE if twice else 0
    ch, sh, rh = engine.tr_svd_host(Ah, dims, ranks, slice_cols=37, precision="fp32", flags=flags)
    cd, sd, rd = engine.tr_svd(A, dims, ranks, precision="fp32")
    torch.cuda.synchronize()
    assert rh["ranks"] == rd["ranks"] == ranks
    k0 = min(dims[0], ranks[0] * ranks[1])
    np.testing.assert_allclose(sh[0][:k0], sd[0][:k0], rtol=0, atol=2e-6 * sd[0][0])
    np.testing.assert_allclose(_cores_host(ch)[0], _cores_host(cd)[0], rtol=0, atol=2e-5)
    a = A.cpu().numpy().astype(np.float64)
    e_h = ttsvd.rel_error(a, ttsvd.tr_full(_cores_host(ch), dims, ranks))
    e_d = ttsvd.rel_error(a, ttsvd.tr_full(_cores_host(cd), dims, ranks))
    assert e_h < 1e-4 and abs(e_h - e_d) < 2e-6


def test_host_stream_rejects_what_it_does_not_cover(engine):
    A = engine.generate([80, 8, 8], [1, 4, 4, 1], seed=3).cpu()
    with pytest.raises(L.TtxError, match="N_1"):
        engine.tt_svd_host(A, [80, 8, 8], [1, 4, 4, 1])
    B = engine.generate([16, 64], [1, 4, 1], seed=3).cpu()
    with pytest.raises(L.TtxError, match="order"):
        engine.tt_svd_host(B, [16, 64], [1, 4, 1])
    C = engine.generate([16, 8, 8], [1, 4, 4, 1], seed=3).cpu()
    with pytest.raises(L.TtxError, match="slice"):
        engine.tt_svd_host(C, [16, 8, 8], [1, 4, 4, 1], slice_cols=-1)
    with pytest.raises(TypeError):
        engine.tt_svd_host(C.cuda(), [16, 8, 8], [1, 4, 4, 1])
    with pytest.raises(L.TtxError, match="fixed ranks"):
        engine.tr_svd_host(C, [16, 8, 8], [2, 2, 4, 2], rank_mode="delta")
    with pytest.raises(L.TtxError, match="r_0 r_1"):
        engine.tr_svd_host(C, [16, 8, 8], [8, 9, 4, 8])
