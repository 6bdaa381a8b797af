"""The headline workload at its full size (order-6 32^6, 4.295 GB; ranks 32, 1e-6 relative noise), in
the bench's arithmetic (BF16, adaptive range check, three-product projection, learned ranks with
ceiling 32 and eps = 1e-4), checked through properties that do not need the oracle to run at this
size:

* the first step's singular values against an FP64 Gram of the 32 x 32^5 unfolding (torch, on the
  device, in column blocks), and their squares summing to ||A||_F^2;
* the cores 1..5 left-orthonormal (G^T G = I), as TT-SVD makes them;
* the relative reconstruction error within eps and near the noise level;
* the same decomposition with fixed ranks and four-product arithmetic agreeing with it."""
import numpy as np
import pytest
import torch

from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L

pytestmark = pytest.mark.gpu

DIMS = [32] * 6
RANKS = [1, 32, 32, 32, 32, 32, 1]


@pytest.fixture(scope="module")
def tensor(engine):
    A = engine.generate(DIMS, RANKS, std=32 ** -0.5, noise_rel=1e-6, seed=1134)
    yield A
    del A
    torch.cuda.empty_cache()


def _gram_fp64(A, m):
    """A_1 A_1^T in FP64 for the m x (numel / m) column-major unfolding, in column blocks."""
    cols = A.view(-1, m)                # row j of this view is column j of A_1
    G = torch.zeros(m, m, dtype=torch.float64, device=A.device)
    step = 1 << 22
    for c0 in range(0, cols.shape[0], step):
        blk = cols[c0:c0 + step].double()
        G += blk.T @ blk
    return G


def test_headline_config_full_size(engine, tensor):
    A = tensor
    flags = L.FLAG_ADAPTIVE | L.FLAG_TAIL_X3
    print("full size: 32^6 learned ranks eps 1e-4, bf16, adaptive, tail x3", flush=True)
    cores, sig, rep = engine.tt_svd(A, DIMS, RANKS, s=10, q=2, precision="bf16", flags=flags,
                                    rank_mode="learned", eps=1e-4)
    torch.cuda.synchronize()
    assert rep["ranks"] == RANKS
    err = engine.rel_error(cores, DIMS, rep["ranks"], A)
    assert err < 2e-5, err
    # step 0 against the FP64 Gram of the unfolding
    G = _gram_fp64(A, 32)
    lam = np.clip(np.linalg.eigvalsh(G.cpu().numpy())[::-1], 0.0, None)
    s_ref = np.sqrt(lam)
    np.testing.assert_allclose(sig[0][:32], s_ref, rtol=0, atol=1e-4 * s_ref[0])
    norm2 = float(torch.trace(G))     # ||A||_F^2
    assert abs(float(np.sum(sig[0][:32].astype(np.float64) ** 2)) - norm2) <= 1e-4 * norm2
    # left-orthonormal cores 1..5: (r_{k-1} N_k) x r_k with orthonormal columns
    for k in range(5):
        g, (rows, cols) = cores[k]
        M = g[: rows * cols].view(cols, rows).T.double()     # column-major rows x cols
        I = M.T @ M
        assert torch.allclose(I, torch.eye(cols, dtype=torch.float64, device=I.device), atol=1e-5), k


def test_headline_arithmetic_agrees_with_four_products(engine, tensor):
    A = tensor
    print("full size: fixed ranks, tail x3 vs x4", flush=True)
    c3, s3, r3 = engine.tt_svd(A, DIMS, RANKS, precision="bf16", flags=L.FLAG_ADAPTIVE | L.FLAG_TAIL_X3)
    e3 = engine.rel_error(c3, DIMS, r3["ranks"], A)
    del c3
    c4, s4, r4 = engine.tt_svd(A, DIMS, RANKS, precision="bf16", flags=L.FLAG_ADAPTIVE)
    e4 = engine.rel_error(c4, DIMS, r4["ranks"], A)
    torch.cuda.synchronize()
    assert r3["ranks"] == r4["ranks"] == RANKS
    assert e3 < 2e-5 and e4 < 2e-5 and abs(e3 - e4) < 5e-6, (e3, e4)
    for k in range(5):
        np.testing.assert_allclose(s3[k][:32], s4[k][:32], rtol=0, atol=1e-4 * s4[k][0])
