"""FP64 CPU restatement of the reference TT-SVD / TR-SVD hot path.

TEST INFRASTRUCTURE ONLY.  This is the checker the HIP engine is compared against; only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it, and the
engine never calls it.

Storage is column-major (Fortran order) throughout, as in the reference: a tensor element
(i1..id) lives at i1 + N1*(i2 + N2*(i3 + ...)).  A TT core G_k is the (r_{k-1}*N_k) x r_k matrix of
the 3-way array (r_{k-1}, N_k, r_k) (ttdec6_singleGPU.cu:194-263, contraction order :141-172).

Functions and the reference code each one restates:

* ``gram_eig_step``      -- longMatrixSVD_Eign_once_device, ttdec6_singleGPU.cu:510-536
                            (Gram A A^T :388-401, syevd :334-385, top-r column reversal :274-290,
                            projection G^T A :424-488)
* ``rsvd_step``          -- basicRandSVD_LU, third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:557-775
* ``cholqr2``            -- replaces basicQR (geqrf + orgqr, rsvd_lu_half.cu:85-159); the Q factor
                            with positive-diagonal R is unique, so both name the same matrix
* ``tt_svd``             -- ttdec_half_lu_device_6, ttdec6_singleGPU.cu:799-958 (order d)
* ``tr_svd``             -- ttdec_half_lu_device_6 (TR), trdec6_singleGPU.cu:827-1004
* ``permute_231``        -- tensorToMode231, trdec6_singleGPU.cu:49-58
* ``permute_312``        -- tensorToMode3_host, trdec6_singleGPU.cu:166-173
* ``tt_full``/``tr_full``-- the contraction chains of calMSE / generalTTTensor / generalTRTensor,
                            ttdec6_singleGPU.cu:134-271, trdec6_singleGPU.cu:175-329
* ``rel_error``          -- calMSE + norm2HH, ttdec6_singleGPU.cu:122-182 (non-destructive here)
* ``rsvd``               -- the basicRandSVD_LU interface: all ell columns of U and SVT
* ``adaptive_range``     -- the randomized range finder with an a-posteriori check on held-out Gaussian
                            test columns (Halko, Martinsson, Tropp 2011, sec. 4.3-4.4); the
                            reference has none (its q and ell are fixed, rsvd_lu_half.cu:557-683), so
                            this is the engine's own rule, pinned by closed forms in the tests
* ``tt_round``/``tt_add``-- TT rounding (Oseledets 2011, Algorithm 2: QR right to left, truncated SVD left
                            to right) and the TT sum whose ranks it recompresses; the reference has no
                            consumer of its cores (they are freed, ttdec6_singleGPU.cu:947-952)
* ``delta_rank``         -- the commented truncation rule,
                            trdecomposition/third_order_GPU/unopt_GPU/trdec_gpu_float_unopt2_rank.cu:92,222-243
"""
import numpy as np

from . import philox

EXACT = "exact"
RSVD = "rsvd"
MAX_ELL = 64          # widest sketch (basis + test columns) of one step


# ----------------------------------------------------------------------------------------------
# small dense pieces
# ----------------------------------------------------------------------------------------------

def gauge(G, *rows_of):
    """Fix the sign freedom of singular vectors: the largest-|.| entry of every column of G is
    made positive, and the matching rows of each array in ``rows_of`` flip with it."""
    G = np.array(G, dtype=np.float64, copy=True)
    idx = np.argmax(np.abs(G), axis=0)
    sgn = np.sign(G[idx, np.arange(G.shape[1])])
    sgn[sgn == 0] = 1.0
    G *= sgn[None, :]
    out = [G]
    for R in rows_of:
        R = np.array(R, dtype=np.float64, copy=True)
        R[: len(sgn)] *= sgn[:, None]
        out.append(R)
    return out if rows_of else G


def cholqr2(Y):
    """Q with orthonormal columns and upper-triangular R (positive diagonal) such that Y = Q R.

    CholeskyQR2 (Fukaya, Nakatsukasa, Yanagisawa, Yamamoto 2014) in FP64; a column whose Gram
    pivot is not positive is handled by Householder QR instead (same Q up to the sign convention,
    which is then fixed to a positive diagonal)."""
    Y = np.asarray(Y, dtype=np.float64)
    try:
        R1 = np.linalg.cholesky(Y.T @ Y).T
        Q1 = np.linalg.solve(R1.T, Y.T).T
        R2 = np.linalg.cholesky(Q1.T @ Q1).T
        Q = np.linalg.solve(R2.T, Q1.T).T
        R = R2 @ R1
    except np.linalg.LinAlgError:
        Q, R = np.linalg.qr(Y)
        s = np.sign(np.diag(R))
        s[s == 0] = 1.0
        Q = Q * s[None, :]
        R = R * s[:, None]
    return Q, R


def qr_pos(Y):
    """Householder QR with the sign fixed so that diag(R) > 0 (the unique thin QR)."""
    Q, R = np.linalg.qr(np.asarray(Y, dtype=np.float64))
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return Q * s[None, :], R * s[:, None]


def eig_desc(W):
    """Symmetric eigen-decomposition with eigenvalues in descending order.

    syevd returns ascending eigenpairs (ttdec6_singleGPU.cu:334-385); the reference then reverses
    the column order (matrixInvertColumn, :274-290).  This returns the reversed form directly."""
    lam, V = np.linalg.eigh(np.asarray(W, dtype=np.float64))
    return lam[::-1].copy(), V[:, ::-1].copy()


# ----------------------------------------------------------------------------------------------
# one decomposition step
# ----------------------------------------------------------------------------------------------

def gram_eig_step(A, r):
    """Exact step: G = top-r eigenvectors of A A^T (descending), A_next = G^T A.

    ttdec6_singleGPU.cu:510-536.  Returns (G m x r, sigma (all min(m,n)) descending, A_next r x n).
    When r exceeds the number of rows the extra columns of G are zero (TR first step with
    r0*r1 > N1)."""
    A = np.asarray(A, dtype=np.float64)
    m = A.shape[0]
    lam, V = eig_desc(A @ A.T)
    lam = np.maximum(lam, 0.0)
    G = np.zeros((m, r))
    kk = min(r, m)
    G[:, :kk] = V[:, :kk]
    G = gauge(G)
    A_next = G.T @ A
    return G, np.sqrt(lam), A_next


def rsvd_tail(A, Q):
    """B = Q^T A (:690-709); W = B B^T, eig, reverse (:717-728); SVT = U_B^T B (:734-739);
    U = Q U_B (:745-763) with the sign gauge.  Returns (U, sigma, SVT)."""
    B = Q.T @ A
    lam, Ub = eig_desc(B @ B.T)
    lam = np.maximum(lam, 0.0)
    U = Q @ Ub
    U, Ub_t = gauge(U, Ub.T)
    SVT = Ub_t @ B
    return U, np.sqrt(lam), SVT


def rsvd_step(A, k, s, q, Omega):
    """Randomized step (EDrSVD), basicRandSVD_LU, rsvd_lu_half.cu:557-775.

    Y = A Omega, Q = qr(Y); q times: Z = A^T Q, Qz = qr(Z), Y = A Qz, Q = qr(Y) (:601-683);
    then rsvd_tail.  Returns (U m x ell, sigma ell, SVT ell x n)."""
    A = np.asarray(A, dtype=np.float64)
    Om = np.asarray(Omega, dtype=np.float64)
    Q, _ = cholqr2(A @ Om)
    for _ in range(q):
        Qz, _ = cholqr2(A.T @ Q)
        Q, _ = cholqr2(A @ Qz)
    return rsvd_tail(A, Q)


def range_resid2(Q, Yt):
    """sum ||(I - Q Q^T) y_t||^2 / sum ||y_t||^2 over the test columns y_t = A omega_t: an unbiased
    estimate of ||(I - Q Q^T) A||_F^2 / ||A||_F^2 for Gaussian omega_t independent of Q."""
    R = Yt - Q @ (Q.T @ Yt)
    y2 = float(np.sum(Yt * Yt))
    return float(np.sum(R * R)) / y2 if y2 > 0 else 0.0


def adaptive_range(A, widths, pt, q, tol2, seed, stream):
    """The engine's adaptive range finder.  For each basis width w in ``widths`` (ascending): sketch
    Y = A Omega[:, :w+pt], Q = orth(Y[:, :w]), check with the test columns Y[:, w:w+pt]; stop at
    the first width that passes.  Then power iterations (Q = orth(A A^T Q)) while the check still
    fails, at most q.  Returns (Q, w, power_iters, rel2, checks)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[1]
    checks = 0
    for w in widths:
        Om = philox.omega(seed, stream, n, w + pt).astype(np.float64)
        Y = A @ Om
        Q, _ = cholqr2(Y[:, :w])
        Yt = Y[:, w:]
        rel2 = range_resid2(Q, Yt)
        checks += 1
        if rel2 <= tol2:
            break
    it = 0
    while it < q and rel2 > tol2:
        Q, _ = cholqr2(A @ (A.T @ Q))
        it += 1
        rel2 = range_resid2(Q, Yt)
        checks += 1
    return Q, w, it, rel2, checks


def adaptive_widths(ell, rank_mode, pred_ell=0):
    """Basis widths the engine tries: the fixed width ell; in delta mode 16, 32, ... up to ell; with a
    learned prediction, the prediction first and then the wider widths of that ladder."""
    if rank_mode == "fixed" and pred_ell <= 0:
        return [ell]
    w0 = min(max(pred_ell, 1), ell) if pred_ell > 0 else min(16, ell)
    out = [w0]
    w = 16
    while w < ell and len(out) < 7:
        if w > w0:
            out.append(w)
        w *= 2
    if out[-1] < ell:
        out.append(ell)
    return out


def step(A, r, s, q, seed, stream, method=RSVD, adaptive=None):
    """One sweep step with the engine's path rule: the exact Gram step when ell = r + s reaches
    min(m, n) or when ``method`` is EXACT; the randomized step otherwise.

    adaptive (optional dict: tol2, rank_mode, pred_ell): the randomized step runs
    ``adaptive_range`` with up to 8 test columns inside the 64-column sketch; the dict receives
    "out" = {ell, power_iters, rel2, checks}."""
    m, n = A.shape
    ell = r + s
    if method == EXACT or ell >= min(m, n):
        G, sig, A_next = gram_eig_step(A, r)
        return G, sig, A_next, EXACT
    ell = min(ell, MAX_ELL)
    pt = min(8, MAX_ELL - ell) if adaptive is not None else 0
    if pt < 2:
        Om = philox.omega(seed, stream, n, ell)
        U, sig, SVT = rsvd_step(A, r, s, q, Om)
        return U[:, :r], sig, SVT[:r], RSVD
    widths = adaptive_widths(ell, adaptive.get("rank_mode", "fixed"), adaptive.get("pred_ell", 0))
    Q, w, it, rel2, checks = adaptive_range(np.asarray(A, dtype=np.float64), widths, pt, q, adaptive["tol2"],
                                            seed, stream)
    U, sig, SVT = rsvd_tail(np.asarray(A, dtype=np.float64), Q)
    adaptive["out"] = dict(ell=w, power_iters=it, rel2=rel2, checks=checks)
    rr = min(r, w)
    return U[:, :rr], sig, SVT[:rr], RSVD


def rsvd(A, k, s, q, seed, stream=0):
    """basicRandSVD_LU's outputs (rsvd_lu_half.cu:557-775): U (m x ell), sigma (ell), SVT (ell x n),
    ell = k + s, with the sketch drawn from the project's Philox stream."""
    A = np.asarray(A, dtype=np.float64)
    Om = philox.omega(seed, stream, A.shape[1], k + s)
    return rsvd_step(A, k, s, q, Om)


# ----------------------------------------------------------------------------------------------
# sweeps
# ----------------------------------------------------------------------------------------------

def delta_rank(lam, total2, delta2, rmax, rel2=None):
    """Rank of the delta rule on one step (Oseledets 2011, TT-SVD; the reference's commented rule
    trdecomposition/third_order_GPU/unopt_GPU/trdec_gpu_float_unopt2_rank.cu:92,222-243 with
    ep = eps / sqrt(d-1)): the smallest r with tail(r) <= delta2, lam = sigma^2 descending, capped at
    rmax and at least 1.

    rel2 None: tail(r) = max(total2 - sum_{i<=r} lam_i, sum_{r<i} lam_i) with total2 = ||A_k||_F^2.
    rel2 given (the basis' relative range residual, 0 for a complete basis):
    tail(r) = sum_{r<i} lam_i + rel2 / (1 - rel2) * sum_i lam_i."""
    lam = np.maximum(np.asarray(lam, dtype=np.float64), 0.0)
    k = len(lam)
    r = k
    cum = np.cumsum(lam)
    allv = float(cum[-1]) if k else 0.0
    outside = 0.0 if rel2 is None else (rel2 / (1.0 - rel2) * allv if rel2 < 1.0 else np.inf)
    for i in range(k):
        if rel2 is None:
            tail = max(total2 - cum[i], float(lam[i + 1:].sum()))
        else:
            tail = float(lam[i + 1:].sum()) + outside
        if tail <= delta2:
            r = i + 1
            break
    return max(1, min(r, rmax))


def tt_svd(A, dims, ranks, s=10, q=2, seed=1234, method=RSVD, rank_mode="fixed", eps=1e-4, adaptive=False,
           infos=None):
    """Order-d TT-SVD sweep (ttdec6_singleGPU.cu:799-958 generalised to order d).

    ranks has d+1 entries with ranks[0] = ranks[d] = 1 (fixed ranks, or ceilings for
    rank_mode="delta").  Returns (cores, sigmas, paths, used_ranks) where core k is the
    (r_{k-1} N_k) x r_k matrix (the last is r_{d-1} x N_d) in FP64."""
    d = len(dims)
    A = np.asarray(A, dtype=np.float64)
    cur = A.reshape(dims[0], -1, order="F")
    norm2 = float(np.sum(A * A))
    delta2 = eps * eps / (d - 1) * norm2
    cores, sigmas, paths = [], [], []
    used = [1]
    total2 = norm2
    for k in range(d - 1):
        m = used[k] * dims[k]
        cur = cur.reshape(m, -1, order="F")
        r = min(ranks[k + 1], cur.shape[0], cur.shape[1])    # min(m, n) at most, as the engine
        ad = None
        if adaptive or rank_mode != "fixed":
            # a quarter of the step's share of the squared error budget, relative to ||A_k||^2
            tol2 = 0.25 * eps * eps / (d - 1) if rank_mode == "fixed" else 0.25 * delta2 / total2
            ad = dict(tol2=tol2, rank_mode=rank_mode)
        if rank_mode == "delta":
            G, sig, nxt, path = step(cur, min(r, min(cur.shape)), s, q, seed, k, method, ad)
            # exact Gram step: the spectrum is complete (rel2 = 0); checked randomized step: its
            # residual estimate; unchecked randomized step: the carried norm
            rel2 = 0.0 if path == EXACT else (ad["out"]["rel2"] if ad and "out" in ad else None)
            rr = delta_rank(sig ** 2, total2, delta2, G.shape[1], rel2)
            G, nxt = G[:, :rr], nxt[:rr]
        else:
            G, sig, nxt, path = step(cur, r, s, q, seed, k, method, ad)
            rr = r
        if infos is not None:
            infos.append(ad.get("out") if ad else None)
        total2 = float(np.sum(sig[:rr] ** 2))
        cores.append(G)
        sigmas.append(sig)
        paths.append(path)
        used.append(rr)
        cur = nxt
    used.append(1)
    cores.append(cur.reshape(used[d - 1], dims[d - 1], order="F"))
    return cores, sigmas, paths, used


def permute_231(T, r0, r1, n):
    """tensorToMode231 (trdec6_singleGPU.cu:49-58): T1 (r0, r1, n) -> T2 (r1, n, r0), column-major;
    T2[b + r1*j + r1*n*a] = T1[a + r0*b + r0*r1*j]."""
    T = np.asarray(T).reshape(-1, order="F")
    return np.ascontiguousarray(
        T.reshape((r0, r1, n), order="F").transpose(1, 2, 0).reshape(-1, order="F"))


def permute_312(T, r1, n, r0):
    """tensorToMode3_host (trdec6_singleGPU.cu:166-173): (r1, n, r0) -> (r0, r1, n)."""
    T = np.asarray(T).reshape(-1, order="F")
    return np.ascontiguousarray(
        T.reshape((r1, n, r0), order="F").transpose(2, 0, 1).reshape(-1, order="F"))


def tr_svd(A, dims, ranks, s=10, q=2, seed=1234, method=RSVD, rank_mode="fixed", eps=1e-4):
    """Order-d TR-SVD sweep, ttdec_half_lu_device_6 (TR), trdec6_singleGPU.cu:827-1004.

    ranks has d+1 entries; ranks[0] = ranks[d] = r0 is the ring rank and the first step keeps
    r0*r1 columns (:865).  G1 is returned as N1 x (r0*r1) with column a + r0*b (not permuted), the
    first remainder is permuted (r0, r1, N') -> (r1, N', r0) (:873), and the last core is
    r_{d-1} x (N_d * r0).

    rank_mode="delta": ranks[1..d-1] are ceilings and every step truncates by the delta rule with
    delta^2 = eps^2 / (d-1) ||A||_F^2 (the reference's commented TR rule,
    trdecomposition/third_order_GPU/unopt_GPU/trdec_gpu_float_unopt2_rank.cu:92,222-243, ep =
    eps / sqrt(d-1)).  The ring rank r0 stays as given: the first step keeps r0 * ceil(R / r0)
    columns for its delta rank R, so r1 = ceil(R / r0) <= ranks[1] and its error stays within delta.
    Returns the ranks used."""
    d = len(dims)
    r0 = ranks[0]
    A = np.asarray(A, dtype=np.float64)
    A1 = A.reshape(dims[0], -1, order="F")
    nrest = A1.shape[1]
    norm2 = float(np.sum(A * A))
    delta2 = eps * eps / (d - 1) * norm2
    delta = rank_mode == "delta"
    # as tt_svd: delta ranks run the range check, a quarter of the step's share of the squared error
    # budget relative to ||A_k||^2, and judge the tail with its residual
    ad = dict(tol2=0.25 * delta2 / norm2, rank_mode=rank_mode) if delta else None
    G1, sig1, A2t, path = step(A1, r0 * ranks[1], s, q, seed, 0, method, ad)
    r1 = ranks[1]
    if delta:
        rel2 = 0.0 if path == EXACT else (ad["out"]["rel2"] if "out" in ad else None)
        R = delta_rank(sig1 ** 2, norm2, delta2, r0 * ranks[1], rel2)
        r1 = min(ranks[1], -(-R // r0))
        G1, A2t = G1[:, :r0 * r1], A2t[:r0 * r1]
    total2 = float(np.sum(sig1[:r0 * r1] ** 2))
    cur = permute_231(A2t.reshape(-1, order="F"), r0, r1, nrest)
    cores, sigmas, paths = [G1], [sig1], [path]
    used = [r0, r1]
    ncols = nrest * r0
    for k in range(1, d - 1):
        m = used[k] * dims[k]
        ncols //= dims[k]
        cur = np.asarray(cur).reshape(m, ncols, order="F")
        if delta:
            rk = min(ranks[k + 1], m, ncols)
            ad = dict(tol2=0.25 * delta2 / total2, rank_mode=rank_mode)
            G, sig, nxt, path = step(cur, rk, s, q, seed, k, method, ad)
            rel2 = 0.0 if path == EXACT else (ad["out"]["rel2"] if "out" in ad else None)
            rr = delta_rank(sig ** 2, total2, delta2, G.shape[1], rel2)
            G, cur = G[:, :rr], nxt[:rr]
        else:
            G, sig, cur, path = step(cur, ranks[k + 1], s, q, seed, k, method)
            rr = ranks[k + 1]
        total2 = float(np.sum(sig[:rr] ** 2))
        cores.append(G)
        sigmas.append(sig)
        paths.append(path)
        used.append(rr)
    used.append(r0)
    cores.append(np.asarray(cur).reshape(used[d - 1], dims[d - 1] * r0, order="F"))
    return cores, sigmas, paths, used


# ----------------------------------------------------------------------------------------------
# reconstruction and error
# ----------------------------------------------------------------------------------------------

def tt_full(cores, dims, ranks):
    """Contract TT cores right to left (generalTTTensor / calMSE order, ttdec6_singleGPU.cu:141-172)."""
    d = len(dims)
    R = np.asarray(cores[-1], dtype=np.float64).reshape(ranks[d - 1], -1, order="F")
    for k in range(d - 2, -1, -1):
        Gk = np.asarray(cores[k], dtype=np.float64).reshape(ranks[k] * dims[k], ranks[k + 1], order="F")
        R = (Gk @ R).reshape(ranks[k], -1, order="F")
    return R.reshape(-1, order="F")


def tr_full(cores, dims, ranks):
    """TR reconstruction as calMSE (TR), trdec6_singleGPU.cu:175-228: chain G2..Gd into
    (r1, N2..Nd, r0), permute to (r0, r1, N'), multiply by G1 (N1 x r0 r1)."""
    d = len(dims)
    r0, r1 = ranks[0], ranks[1]
    R = np.asarray(cores[-1], dtype=np.float64).reshape(ranks[d - 1], -1, order="F")
    for k in range(d - 2, 0, -1):
        Gk = np.asarray(cores[k], dtype=np.float64).reshape(ranks[k] * dims[k], ranks[k + 1], order="F")
        R = (Gk @ R).reshape(ranks[k], -1, order="F")
    nrest = int(np.prod(dims[1:]))
    mid = permute_312(R.reshape(-1, order="F"), r1, nrest, r0).reshape(r0 * r1, nrest, order="F")
    G1 = np.asarray(cores[0], dtype=np.float64).reshape(dims[0], r0 * r1, order="F")
    return (G1 @ mid).reshape(-1, order="F")


def tt_add(ca, cb, dims, ra, rb):
    """Cores of the sum of two TTs (Oseledets 2011, sec. 4.1): block-diagonal cores of ranks ra + rb, the
    first core [A1 B1], the last [Ad; Bd]."""
    d = len(dims)
    rs = [1] + [ra[k] + rb[k] for k in range(1, d)] + [1]
    out = []
    for k in range(d):
        A = np.asarray(ca[k], dtype=np.float64).reshape(ra[k], dims[k], ra[k + 1], order="F")
        B = np.asarray(cb[k], dtype=np.float64).reshape(rb[k], dims[k], rb[k + 1], order="F")
        C = np.zeros((rs[k], dims[k], rs[k + 1]))
        if k == 0:
            C[:, :, :ra[1]] = A
            C[:, :, ra[1]:] = B
        elif k == d - 1:
            C[:ra[k], :, :] = A
            C[ra[k]:, :, :] = B
        else:
            C[:ra[k], :, :ra[k + 1]] = A
            C[ra[k]:, :, ra[k + 1]:] = B
        out.append(C.reshape(rs[k] * dims[k], rs[k + 1], order="F"))
    return out, rs


def tt_round(cores, dims, ranks, eps, max_ranks=None):
    """TT rounding in FP64 (Oseledets 2011, Algorithm 2): right to left, core k (r_{k-1} x N_k r_k) = R^T Q^T
    from the QR of its transpose, core k <- Q^T and R^T moves left; then left to right, the SVD of core k's
    (r_{k-1} N_k) x r_k unfolding truncated by the delta rule, delta^2 = eps^2 ||A||^2 / (d - 1) (the
    tail sum of squared singular values at most delta^2), core k <- U_r and S V^T moves right.
    Returns (cores, ranks, sigma per bond)."""
    d = len(dims)
    r = list(ranks)
    G = [np.asarray(c, dtype=np.float64).reshape(r[k] * dims[k], r[k + 1], order="F").copy()
         for k, c in enumerate(cores)]
    for k in range(d - 1, 0, -1):
        M = G[k].reshape(r[k], dims[k] * r[k + 1], order="F")
        Q, R = np.linalg.qr(M.T)
        rk = Q.shape[1]
        G[k] = Q.T.reshape(rk * dims[k], r[k + 1], order="F")
        G[k - 1] = G[k - 1] @ R.T
        r[k] = rk
    norm2 = float(np.sum(G[0] ** 2))
    delta2 = eps * eps * norm2 / (d - 1)
    sig = []
    for k in range(d - 1):
        U, S, Vt = np.linalg.svd(G[k], full_matrices=False)
        cap = r[k + 1] if max_ranks is None else max(1, min(max_ranks[k + 1], r[k + 1]))
        rn = delta_rank(S ** 2, float(np.sum(S ** 2)), delta2, min(cap, len(S)), rel2=0.0)
        sig.append(S)
        G[k] = U[:, :rn]
        nxt = (S[:rn, None] * Vt[:rn]) @ G[k + 1].reshape(r[k + 1], dims[k + 1] * r[k + 2], order="F")
        G[k + 1] = nxt.reshape(rn * dims[k + 1], r[k + 2], order="F")
        r[k + 1] = rn
    return G, r, sig


def rel_error(A, Ahat):
    """||A - Ahat||_F / ||A||_F in FP64 (calMSE, ttdec6_singleGPU.cu:134-182, without overwriting A)."""
    A = np.asarray(A, dtype=np.float64).reshape(-1)
    Ahat = np.asarray(Ahat, dtype=np.float64).reshape(-1)
    return float(np.linalg.norm(A - Ahat) / np.linalg.norm(A))


# ----------------------------------------------------------------------------------------------
# synthetic inputs (the BASELINE distributions, drawn from Philox)
# ----------------------------------------------------------------------------------------------

DATA_SEED = 1134      # mirrors srand(1134), ttdec6_singleGPU.cu:986
OMEGA_SEED = 1234     # mirrors the cuRAND seed, rsvd_lu_half.cu:586
NOISE_STREAM = 0xFFFF


def tt_core_shapes(dims, ranks):
    return [(ranks[k] * dims[k], ranks[k + 1]) for k in range(len(dims))]


def make_tt_cores(dims, ranks, std=1.0, seed=DATA_SEED):
    """Core k drawn iid N(0, std^2) from Philox stream 0x100 + k (float32)."""
    cores = []
    for k, (a, b) in enumerate(tt_core_shapes(dims, ranks)):
        z = philox.normal_flat(seed, 0x100 + k, a * b) * np.float32(std)
        cores.append(z.astype(np.float32).reshape(a, b, order="F"))
    return cores


def make_tr_cores(dims, ranks, std=1.0, seed=DATA_SEED):
    """Ring cores in the reference's storage: G1 is N1 x (r0 r1), G2..G_{d-1} are
    (r_{k-1} N_k) x r_k and G_d is r_{d-1} x (N_d r0); iid N(0, std^2), Philox stream 0x200 + k."""
    d = len(dims)
    shapes = [(dims[0], ranks[0] * ranks[1])]
    shapes += [(ranks[k] * dims[k], ranks[k + 1]) for k in range(1, d - 1)]
    shapes += [(ranks[d - 1], dims[d - 1] * ranks[0])]
    cores = []
    for k, (a, b) in enumerate(shapes):
        z = philox.normal_flat(seed, 0x200 + k, a * b) * np.float32(std)
        cores.append(z.astype(np.float32).reshape(a, b, order="F"))
    return cores


def add_noise(A, rel, seed=DATA_SEED):
    """A + rel * rms(A) * N(0,1) (float32), noise from Philox stream NOISE_STREAM."""
    A = np.asarray(A, dtype=np.float32).reshape(-1)
    if rel == 0:
        return A.copy()
    rms = np.float32(np.sqrt(np.mean(A.astype(np.float64) ** 2)))
    z = philox.normal_flat(seed, NOISE_STREAM, A.size)
    return (A + np.float32(rel) * rms * z).astype(np.float32)
