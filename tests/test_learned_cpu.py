"""Rank model on the CPU: the oracle's backward pass against central differences, the training-data
features and probe pivots against the oracle's independent statement, label closed forms, and the
learned-rank workload generator.  CPU only."""
import numpy as np

from oracle import mlp as om
from tensor_train_ring_decompositions_hpc_gpu_amd import learned as Lr


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(om.PARAMS) * scale


def test_oracle_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    X = rng.standard_normal((7, om.IN))
    T = rng.standard_normal((7, om.OUT))
    loss, g = om.loss_and_grad(p, X, T, rounding=False)
    idx = rng.choice(om.PARAMS, 60, replace=False)
    # every block: W1, b1, W2, b2, W3, b3
    idx = np.concatenate([idx, [0, om.IN * om.HID + 3, om.IN * om.HID + om.HID + 5, om.PARAMS - om.OUT - 7,
                                om.PARAMS - 1]])
    h = 1e-6
    for i in idx:
        e = np.zeros(om.PARAMS)
        e[i] = h
        fd = (om.loss_and_grad(p + e, X, T, False)[0] - om.loss_and_grad(p - e, X, T, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_rounded_forward_is_binary16_operands():
    # with binary16-exact inputs and weights the rounded and plain networks agree up to the rounding
    # of the hidden activations
    rng = np.random.default_rng(3)
    p = _params(4).astype(np.float16).astype(np.float64)
    X = rng.standard_normal((5, om.IN)).astype(np.float16).astype(np.float64)
    y16, _ = om.forward(p, X, True)
    y64, _ = om.forward(p, X, False)
    assert np.max(np.abs(y16 - y64)) < 2e-2 * np.max(np.abs(y64))
    assert not np.array_equal(y16, y64)


def test_training_features_match_oracle():
    rng = np.random.default_rng(5)
    piv = 10.0 ** rng.uniform(-12, 2, (6, 64))
    piv[2, 40:] = 0.0
    tol = 10.0 ** rng.uniform(-8, -3, 6)
    m = rng.integers(64, 4096, 6)
    n = rng.integers(1 << 12, 1 << 24, 6)
    a = Lr.features(piv, tol, m, n)
    for b in range(6):
        assert np.allclose(a[b], om.features(piv[b], tol[b], m[b], n[b]), rtol=0, atol=1e-12)


def test_training_pivots_match_oracle_cholesky():
    rng = np.random.default_rng(6)
    s = Lr.sample_spectra(rng, 3)
    for b in range(3):
        Y = s[b][:, None] * rng.standard_normal((Lr.DIM, 64))
        G = Y.T @ Y
        p1 = Lr.shifted_pivots(G[None], np.array([500.0]))[0]
        p2 = om.shifted_cholesky_pivots(G, 500.0)
        assert np.allclose(p1, p2, rtol=1e-8, atol=1e-14 * np.trace(G))


def test_label_closed_forms():
    # delta rank: sigma = (1, 1, 0.1, 0.01), budget 1.5e-4 of 2.0101: tail after 2 = 0.0101 > 3.0e-4,
    # after 3 = 1e-4 <= 3.0e-4 -> rank 3
    s = np.zeros((1, Lr.DIM))
    s[0, :4] = [1.0, 1.0, 0.1, 0.01]
    assert Lr.delta_rank(s, np.array([1.5e-4]))[0] == 3
    # an exact rank-L spectrum is captured by a width-L sketch: the minimal width is at most L
    rng = np.random.default_rng(7)
    for L in (3, 17, 40):
        s = np.zeros((1, Lr.DIM))
        s[0, :L] = np.linspace(1.0, 0.5, L)
        w = Lr.min_width(s, np.array([1e-8]), rng)[0]
        assert w == L
    # more than 56 equal values cannot pass: 57
    s = np.zeros((1, Lr.DIM))
    s[0, :80] = 1.0
    assert Lr.min_width(s, np.array([1e-6]), rng)[0] == Lr.MAX_W + 1


def test_spectral_tt_first_unfolding_spectrum():
    dims, alphas = [6, 7, 5, 6], [1.5, 0.8, 2.0]
    A, r = Lr.spectral_tt(dims, 5, alphas, seed=3)
    A1 = A.astype(np.float64).reshape(dims[0], -1, order="F")
    sv = np.linalg.svd(A1, compute_uv=False)
    assert r == [1, 5, 5, 5, 1]
    assert sv.shape[0] == 6 and sv[5] < 1e-5 * sv[0]
    assert np.all(np.diff(sv[:5]) <= 1e-6)
    assert np.isfinite(A).all()
