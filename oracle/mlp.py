"""FP64 CPU restatement of the rank model (TTX_RANK_LEARNED): forward, loss gradient, input features
and the probe.

TEST INFRASTRUCTURE ONLY: the checker the HIP kernels (csrc/ttx_mlp.hip) and the probe
(csrc/ttx_engine.cpp ``probe``) are compared against; only ``tests/`` may import it.

The reference has no rank model (its ranks are fixed inputs, ttdec6_singleGPU.cu:977); this file
states the network the engine implements:

  Z1 = r16(X) r16(W1) + b1, H1 = GELU(Z1), Z2 = r16(H1) r16(W2) + b2, H2 = GELU(Z2),
  Y = r16(H2) r16(W3) + b3,   loss = mean((Y - T)^2)

with r16 = rounding to IEEE binary16 (the operands of an FP16 MFMA) and GELU(x) = x Phi(x).  The
backward pass rounds the weights and activations it multiplies (r16(H)^T dZ, dZ r16(W)^T) and keeps
the gradient signals dZ in full precision; ``rounding=False`` gives the plain FP64 network, whose
gradient the tests pin by central differences.
"""
import numpy as np
from scipy.special import erf

from . import philox

IN, HID, OUT = 40, 128, 2
PARAMS = IN * HID + HID + HID * HID + HID + HID * OUT + OUT


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + IN * HID].reshape(IN, HID); o += IN * HID
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    W3 = p[o:o + HID * OUT].reshape(HID, OUT); o += HID * OUT
    b3 = p[o:o + OUT]
    return W1, b1, W2, b2, W3, b3


def _r(x, rounding):
    return np.asarray(x, dtype=np.float32).astype(np.float16).astype(np.float64) if rounding else np.asarray(x, np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def forward(p, X, rounding=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    Z1 = _r(X, rounding) @ _r(W1, rounding) + b1
    H1 = gelu(Z1)
    Z2 = _r(H1, rounding) @ _r(W2, rounding) + b2
    H2 = gelu(Z2)
    Y = _r(H2, rounding) @ _r(W3, rounding) + b3
    return Y, (X, Z1, H1, Z2, H2)


def loss_and_grad(p, X, T, rounding=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Y, (X, Z1, H1, Z2, H2) = forward(p, X, rounding)
    T = np.asarray(T, dtype=np.float64)
    D = Y - T
    loss = float(np.mean(D * D))
    dY = 2.0 * D / D.size
    gW3 = _r(H2, rounding).T @ dY
    gb3 = dY.sum(0)
    dZ2 = (dY @ _r(W3, rounding).T) * gelu_d(Z2)
    gW2 = _r(H1, rounding).T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ _r(W2, rounding).T) * gelu_d(Z1)
    gW1 = _r(X, rounding).T @ dZ1
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def features(pivots, rel_tol, m, n):
    """k_mlp_features: 32 log-buckets of the probe pivots, log10 of the relative budget, log2 m,
    log2 n and a constant."""
    d = np.asarray(pivots, dtype=np.float64)
    kp = len(d)
    D = d.sum()
    D = D if D > 0 else 1.0
    x = np.zeros(IN)
    for b in range(32):
        if 2 * b + 1 < kp:
            x[b] = np.log10(max(d[2 * b] + d[2 * b + 1], 1e-30 * D) / D) / 10.0
    x[32] = np.log10(rel_tol) / 10.0
    x[33] = np.log2(m) / 16.0
    x[34] = np.log2(n) / 32.0
    x[35] = 1.0
    return x


def shifted_cholesky_pivots(G, m):
    """Right-looking Cholesky of G + s I with the shifted-CholeskyQR shift s = 11 (m k + k (k+1)) u
    trace(G) (Fukaya et al. 2020), u = 2^-53; pivots not above (max diag + s) 1e-28 count as zero and
    leave their row out (k_chol_rinv)."""
    a = np.array(G, dtype=np.float64)
    k = a.shape[0]
    tr = np.trace(a)
    s = 11.0 * (m * k + k * (k + 1)) * 1.1102230246251565e-16 * tr
    tiny = (np.max(np.diag(a)) + s) * 1e-28
    piv = np.zeros(k)
    for i in range(k):
        d = a[i, i] + s
        if d > tiny:
            piv[i] = d
            r = a[i, i + 1:] / np.sqrt(d)
            a[i + 1:, i + 1:] -= np.outer(r, r)
        else:
            a[i, i + 1:] = 0.0
    return piv


def probe_columns(n):
    """columns of the probe: the first max(min(n, 512), n / 16)"""
    return min(n, max(min(n, 512), n // 16))


def probe_pivots(A, seed, stream):
    """The engine's probe of an m x n unfolding: Y = A[:, :n_p] Omega (64 columns from Philox stream
    ``stream | 0x10000``), pivots of the shifted Cholesky of Y^T Y."""
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    npc = probe_columns(n)
    Om = philox.omega(seed, stream | 0x10000, npc, 64).astype(np.float64)
    Y = A[:, :npc] @ Om
    return shifted_cholesky_pivots(Y.T @ Y, m)
