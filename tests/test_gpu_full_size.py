"""The headline workload at its full size (order-6 32^6, 4.295 GB; ranks 32, 1e-6 relative noise), in
the bench's arithmetic (BF16, adaptive range check, three-product projection, learned ranks with
ceiling 32 and eps = 1e-4), checked through properties that do not need the oracle to run at this
size:

* the first step's squared singular values against the eigenvalues of an FP64 Gram of the 32 x 32^5
  unfolding (torch, on the device, in column blocks), and their sum against ||A||_F^2;
* steps 1..4's singular values against the FP64 singular values (torch on the device: the Gram of the
  smaller side, or for 32^3 x 32^3 a randomized SVD with 16 extra columns and two power iterations) of
  the tensor's own unfolding 32^(k+1) x 32^(5-k): the tensor has TT ranks 32
  plus noise at 1e-6 of ||A||, so the remainder each step sees has the unfolding's top 32 singular
  values up to that noise;
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


def _top_sv_fp64(X64, m, p=48, q=2, seed=5):
    """Largest singular values of the m x (numel / m) column-major unfolding in FP64: the Gram of its
    smaller side when that side is at most 1024, else a randomized SVD (Gaussian start, QR)."""
    X = X64.view(-1, m)                 # n x m: row j is column j of the unfolding
    n = X.shape[0]
    if min(m, n) <= 1024:
        G = (X.T @ X) if m <= n else (X @ X.T)
        lam = np.linalg.eigvalsh(G.cpu().numpy())[::-1]
        return np.sqrt(np.clip(lam, 0.0, None))
    g = torch.Generator(device=X.device).manual_seed(seed)
    Om = torch.randn(n, p, dtype=torch.float64, device=X.device, generator=g)
    Q = torch.linalg.qr(X.T @ Om).Q
    for _ in range(q):
        Qz = torch.linalg.qr(X @ Q).Q
        Q = torch.linalg.qr(X.T @ Qz).Q
    Bt = X @ Q                          # (Q^T A_k)^T, n x p
    lam = torch.linalg.eigvalsh(Bt.T @ Bt).flip(0).clamp_min(0.0)
    return np.sqrt(lam.cpu().numpy())


def _mid_sq(A):
    """sum over the entries of mid^2, mid = bf16(x - bf16(x)) (torch rounds to nearest even, as the
    device conversion does), in FP64 on the device, in blocks"""
    flat = A.view(-1)
    tot = 0.0
    step = 1 << 26
    for c0 in range(0, flat.numel(), step):
        x = flat[c0:c0 + step]
        r = x - x.to(torch.bfloat16).float()
        tot += float((r.to(torch.bfloat16).double() ** 2).sum())
    return tot


def _norm2(A):
    flat = A.view(-1)
    step = 1 << 26
    return sum(float((flat[c0:c0 + step].double() ** 2).sum()) for c0 in range(0, flat.numel(), step))


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
    # step 0 takes sigma = sqrt(eigenvalue) of the fused first pass's Gram, whose three bf16 products
    # drop mid * mid: on the diagonal a sum of squares, so the eigenvalues are low by about 3e-6 of
    # ||A||^2; sigma_i then carries that error divided by 2 sigma_i, so the bound is on sigma^2
    s0 = sig[0][:32].astype(np.float64)
    dev_0 = float(np.max(np.abs(s0 ** 2 - s_ref[:32] ** 2)) / s_ref[0] ** 2)
    print("step 0: max |sigma^2 - sigma_fp64^2| / sigma_1^2 = %.2e" % dev_0, flush=True)
    assert dev_0 < 1e-5, dev_0
    norm2 = float(torch.trace(G))     # ||A||_F^2
    bias = (float(np.sum(s0 ** 2)) - norm2) / norm2
    # the trace of that Gram is sum (hi^2 + 2 hi mid) over the entries (x = hi + mid + lo, bf16 terms
    # rounded to nearest even): low by sum mid^2 / sum x^2, which the tensor itself fixes
    pred = -_mid_sq(A) / norm2
    print("step 0: (sum sigma^2 - ||A||^2) / ||A||^2 = %.3e, dropped mid * mid predicts %.3e" % (bias, pred),
          flush=True)
    assert abs(bias - pred) <= 0.2 * abs(pred) + 1e-7, (bias, pred)
    del G
    # steps 1..4 against an FP64 randomized SVD of the unfoldings
    X64 = A.double()
    for k in range(1, 5):
        s_k = _top_sv_fp64(X64, 32 ** (k + 1))[:32]
        dev_k = float(np.max(np.abs(sig[k][:32].astype(np.float64) - s_k)) / s_k[0])
        print("step %d: max |sigma - sigma_fp64| / sigma_1 = %.2e" % (k, dev_k), flush=True)
        assert dev_k < 2e-6, (k, dev_k)
    del X64
    torch.cuda.empty_cache()
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
    # with four products the first pass's Gram keeps mid * mid: its trace is unbiased (the three-product
    # trace is low by about 2.7e-6 of ||A||^2, checked above)
    norm2 = _norm2(A)
    bias4 = (float(np.sum(s4[0][:32].astype(np.float64) ** 2)) - norm2) / norm2
    print("step 0, four products: (sum sigma^2 - ||A||^2) / ||A||^2 = %.3e" % bias4, flush=True)
    assert abs(bias4) <= 3e-7, bias4
    for k in range(5):
        np.testing.assert_allclose(s3[k][:32], s4[k][:32], rtol=0, atol=1e-4 * s4[k][0])
