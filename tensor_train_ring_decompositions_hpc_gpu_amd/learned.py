"""Training data and training loop of the rank model (TTX_RANK_LEARNED), and the synthetic
learned-rank workload (order-5 tensors whose bonds carry singular values i^-alpha).

The model maps the probe of one unfolding (32 log-buckets of the shifted Cholesky pivots of a
64-column Gaussian sketch) plus the relative error budget and log2 m, log2 n to (rank / 64, sketch
width / 64).  Labels come from the exact truncated method on a known spectrum:

* rank: the delta rule, the smallest r with sum_{i>r} s_i^2 <= budget * sum s_i^2;
* width: the smallest basis width w whose Gaussian sketch leaves a range residual
  ||(I - Q_w Q_w^T) A||_F^2 / ||A||_F^2 <= budget / 4 (the tolerance of the engine's a-posteriori
  check); 57 when no width up to 56 does.

Both are functions of the spectrum and of Gaussian draws only: for A = U diag(s) V^T and Gaussian
Omega, A Omega = U diag(s) (V^T Omega) with V^T Omega again Gaussian, so a sample needs the spectrum
and two Gaussian matrices (probe and sketch), never the matrix itself.

Training runs on the GPU through the engine's own forward / backward / Adam kernels
(``Engine.mlp_train_step``); numpy only generates the samples.

Usage: python -m tensor_train_ring_decompositions_hpc_gpu_amd.learned [--samples 100000] [--out PATH]
"""
import argparse
import os
import sys
import time

import numpy as np

IN, HID, OUT = 40, 128, 2
PROBE_K = 64
MAX_W = 56
DIM = 96             # simulated spectrum length (exact rank plus noise floor)
MARGIN = 2           # columns the engine adds to the predicted width


# ------------------------------------------------------------------------------------------------
# features (the numpy statement of k_mlp_features) and the batched probe simulation
# ------------------------------------------------------------------------------------------------
def features(pivots, rel_tol, m, n):
    """pivots: (B, 64); rel_tol, m, n: (B,).  Returns (B, IN) float64."""
    d = np.asarray(pivots, dtype=np.float64)
    B = d.shape[0]
    D = d.sum(1)
    D = np.where(D > 0, D, 1.0)
    x = np.zeros((B, IN))
    pair = d[:, 0::2] + d[:, 1::2]                       # (B, 32)
    x[:, :32] = np.log10(np.maximum(pair, 1e-30 * D[:, None]) / D[:, None]) / 10.0
    x[:, 32] = np.log10(rel_tol) / 10.0
    x[:, 33] = np.log2(m) / 16.0
    x[:, 34] = np.log2(n) / 32.0
    x[:, 35] = 1.0
    return x


def shifted_pivots(G, m):
    """Pivots R_ii^2 of the Cholesky of G + s I, s = 11 (m k + k (k+1)) u trace(G) (the shifted
    CholeskyQR shift of the engine's probe), batched over the leading axis."""
    k = G.shape[-1]
    tr = np.trace(G, axis1=1, axis2=2)
    s = 11.0 * (m * k + k * (k + 1)) * 1.1102230246251565e-16 * tr
    L = np.linalg.cholesky(G + s[:, None, None] * np.eye(k)[None])
    return np.diagonal(L, axis1=1, axis2=2) ** 2


def sample_spectra(rng, count):
    alpha = rng.uniform(0.3, 3.3, count)
    L = rng.integers(2, 65, count)
    i = np.arange(1, DIM + 1)[None, :]
    s = i ** (-alpha[:, None])
    floor = np.where(rng.random(count) < 0.5, 10.0 ** rng.uniform(-8, -4, count), 0.0)
    s = np.where(i <= L[:, None], s, floor[:, None])
    return s


def delta_rank(s, budget):
    """smallest r with sum_{i>r} s_i^2 <= budget sum s_i^2 (capped at 64)"""
    e = s * s
    tot = e.sum(1)
    tail = tot[:, None] - np.cumsum(e, 1)            # tail after keeping i+1 values
    ok = tail <= budget[:, None] * tot[:, None]
    r = np.argmax(ok, 1) + 1
    r = np.where(ok.any(1), r, DIM)
    return np.minimum(r, 64)


def min_width(s, budget, rng):
    """smallest w <= 56 whose Gaussian sketch passes the range tolerance budget / 4; 57 otherwise"""
    B = s.shape[0]
    G = rng.standard_normal((B, DIM, MAX_W))
    Q, _ = np.linalg.qr(s[:, :, None] * G)                      # (B, DIM, MAX_W), prefix property
    # captured energy of column w: ||Q_w^T diag(s)||^2 = sum_i Q_iw^2 s_i^2
    cap = np.einsum("biw,bi->bw", Q * Q, s * s)
    tot = (s * s).sum(1)
    resid = 1.0 - np.cumsum(cap, 1) / tot[:, None]
    ok = resid <= 0.25 * budget[:, None]
    w = np.argmax(ok, 1) + 1
    return np.where(ok.any(1), w, MAX_W + 1)


def _chunk(args):
    seed, index, b = args
    rng = np.random.default_rng([seed, index])
    s = sample_spectra(rng, b)
    m = np.round(2.0 ** rng.uniform(6, 12, b))
    n = np.round(2.0 ** rng.uniform(12, 26, b))
    eps = 10.0 ** rng.uniform(-4, -1.5, b)
    d = rng.integers(3, 9, b)
    budget = eps * eps / (d - 1)
    Gp = rng.standard_normal((b, DIM, PROBE_K))
    Y = s[:, :, None] * Gp
    piv = shifted_pivots(np.einsum("bik,bil->bkl", Y, Y), m)
    X = features(piv, budget, m, n)
    r = delta_rank(s, budget)
    w = min_width(s, budget, rng)
    return X.astype(np.float32), np.stack([r / 64.0, w / 64.0], 1).astype(np.float32), r, w


def make_dataset(count, seed=2024, chunk=1024, workers=None):
    """(X, T, info) with X (count, IN) float32 features, T (count, OUT) float32 targets; chunks are
    seeded by (seed, chunk index), so the set does not depend on the number of workers"""
    jobs = [(seed, i, min(chunk, count - i * chunk)) for i in range((count + chunk - 1) // chunk)]
    workers = workers or min(8, os.cpu_count() or 1)
    if workers > 1 and len(jobs) > 1:
        from multiprocessing import get_context
        with get_context("spawn").Pool(workers) as pool:
            parts = pool.map(_chunk, jobs)
    else:
        parts = [_chunk(j) for j in jobs]
    X = np.concatenate([p[0] for p in parts])
    T = np.concatenate([p[1] for p in parts])
    return X, T, {"rank": np.concatenate([p[2] for p in parts]), "width": np.concatenate([p[3] for p in parts])}


# ------------------------------------------------------------------------------------------------
# training on the GPU
# ------------------------------------------------------------------------------------------------
def train(eng, X, T, epochs=60, batch=1024, lr=3e-3, seed=7, log=print):
    import torch
    Xd = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    Td = torch.from_numpy(np.ascontiguousarray(T)).cuda()
    eng.mlp_init(seed)
    rng = np.random.default_rng(seed)
    N = X.shape[0]
    steps = 0
    total = epochs * (N // batch)
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(N)).cuda()
        Xp = Xd[perm].contiguous()
        Tp = Td[perm].contiguous()
        losses = []
        for i in range(N // batch):
            # cosine decay of the learning rate
            lr_t = lr * 0.5 * (1.0 + np.cos(np.pi * steps / total))
            want = (i == N // batch - 1)
            l = eng.mlp_train_step(Xp[i * batch:(i + 1) * batch], Tp[i * batch:(i + 1) * batch], max(lr_t, 1e-5),
                                   want_loss=want)
            if want:
                losses.append(l)
            steps += 1
        if log and (ep % 10 == 0 or ep == epochs - 1):
            log("epoch %d  loss %.3e" % (ep, losses[-1]))
    return eng.mlp_get_params()


def evaluate(eng, X, info):
    """rank-hit rate, mean |rank error|, and the fraction whose predicted width (+ margin) covers the
    labelled minimal width"""
    import torch
    Xd = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    Y = np.concatenate([eng.mlp_forward(Xd[i:i + 4096]).cpu().numpy() for i in range(0, X.shape[0], 4096)])
    Y = Y.astype(np.float64)
    r_hat = np.clip(np.round(64 * Y[:, 0]), 1, 64)
    w_hat = np.clip(np.ceil(64 * Y[:, 1]) + MARGIN, 1, MAX_W)
    r, w = info["rank"], info["width"]
    w0 = np.ceil(64 * Y[:, 1])
    return {"rank_hit": float(np.mean(r_hat == r)), "rank_within_2": float(np.mean(np.abs(r_hat - r) <= 2)),
            "rank_mae": float(np.mean(np.abs(r_hat - r))),
            "width_cover": float(np.mean(w_hat >= np.minimum(w, MAX_W))),
            "width_excess": float(np.mean(np.maximum(w_hat - w, 0))),
            "cover_by_margin": {mg: float(np.mean(np.clip(w0 + mg, 1, MAX_W) >= np.minimum(w, MAX_W)))
                                for mg in (0, 2, 4, 6, 8)}}


# ------------------------------------------------------------------------------------------------
# the learned-rank workload: order-d tensors whose bond k carries weights i^-alpha_k
# ------------------------------------------------------------------------------------------------
def spectral_tt(dims, rmax, alphas, seed):
    """A = U_1 D_1 U_2 D_2 ... D_{d-1} U_d (column-major, float32) with seeded orthonormal cores
    U_k ((r_{k-1} N_k) x r_k orthonormal columns; the last (r_{d-1} x N_d) orthonormal rows) and bond
    weights D_k = diag(i^-alpha_k), r_k = min(rmax, prod N_1..N_k, prod N_{k+1}..N_d).  Each
    unfolding's spectrum follows its bond's weights as mixed by the weights of the other bonds (no
    tensor has prescribed spectra at every bond in general)."""
    rng = np.random.default_rng(seed)
    d = len(dims)
    r = [1]
    for k in range(1, d):
        r.append(int(min(rmax, np.prod(dims[:k]), np.prod(dims[k:]))))
    r.append(1)
    cores = []
    for k in range(d):
        rows, cols = r[k] * dims[k], r[k + 1]
        if k == d - 1:
            M = np.linalg.qr(rng.standard_normal((dims[k], r[k])))[0].T        # r_{d-1} x N_d, orthonormal rows
        else:
            M = np.linalg.qr(rng.standard_normal((rows, cols)))[0]
        cores.append(M)
    R = cores[-1].astype(np.float64)
    for k in range(d - 2, -1, -1):
        w = np.arange(1, r[k + 1] + 1, dtype=np.float64) ** (-alphas[k])
        R = cores[k] @ (w[:, None] * R)                                        # (r_k N_k) x rest
        R = R.reshape(r[k], -1, order="F")
    return R.reshape(-1, order="F").astype(np.float32), r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=100000)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "rank_mlp.f32"))
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args()
    from .api import Engine
    t0 = time.time()
    X, T, info = make_dataset(args.samples, seed=args.seed)
    Xv, Tv, info_v = make_dataset(max(2000, args.samples // 10), seed=args.seed + 1)
    print("data: %d train / %d validation samples in %.1f s" % (len(X), len(Xv), time.time() - t0), flush=True)
    eng = Engine()
    t0 = time.time()
    params = train(eng, X, T, epochs=args.epochs, log=lambda s: print(s, flush=True))
    print("training: %.1f s" % (time.time() - t0), flush=True)
    params.astype("<f4").tofile(args.out)
    print("wrote", args.out, flush=True)
    print("validation:", evaluate(eng, Xv, info_v), flush=True)


if __name__ == "__main__":
    sys.exit(main())
