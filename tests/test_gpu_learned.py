"""Rank model on the GPU: forward, gradient and features against the FP64 oracle (oracle/mlp.py),
the probe against the oracle's probe, training on the GPU, and what the trained network gains over
the adaptive range finder without it.

Tolerances: the HIP network and the oracle round the same operands to binary16; they differ by FP32
versus FP64 accumulation, which can move an activation across a binary16 rounding boundary, so
outputs agree to 2e-3 relative and gradients to 1e-2 relative (norm per parameter block)."""
import os

import numpy as np
import pytest

from oracle import mlp as om
from oracle import ttsvd

pytestmark = pytest.mark.gpu


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float32))).cuda()


def _params(seed, scale=0.15):
    return (np.random.default_rng(seed).standard_normal(om.PARAMS) * scale).astype(np.float32)


BLOCKS = [("W1", 0, om.IN * om.HID), ("b1", om.IN * om.HID, om.HID),
          ("W2", om.IN * om.HID + om.HID, om.HID * om.HID),
          ("b2", om.IN * om.HID + om.HID + om.HID * om.HID, om.HID),
          ("W3", om.PARAMS - om.OUT - om.HID * om.OUT, om.HID * om.OUT), ("b3", om.PARAMS - om.OUT, om.OUT)]


def test_mlp_forward_matches_oracle(engine):
    p = _params(1)
    engine.mlp_set_params(p)
    assert np.array_equal(engine.mlp_get_params(), p)
    X = np.random.default_rng(2).standard_normal((300, om.IN)).astype(np.float32)
    Y = engine.mlp_forward(_dev(X)).cpu().numpy()
    ref, _ = om.forward(p.astype(np.float64), X, rounding=True)
    assert np.max(np.abs(Y - ref)) <= 2e-3 * np.max(np.abs(ref))


@pytest.mark.parametrize("batch", [1, 37, 1024])
def test_mlp_gradient_matches_oracle(engine, batch):
    p = _params(3)
    engine.mlp_set_params(p)
    rng = np.random.default_rng(batch)
    X = rng.standard_normal((batch, om.IN)).astype(np.float32)
    T = rng.standard_normal((batch, om.OUT)).astype(np.float32)
    g, loss = engine.mlp_gradient(_dev(X), _dev(T))
    g = g.cpu().numpy().astype(np.float64)
    loss_ref, g_ref = om.loss_and_grad(p.astype(np.float64), X, T, rounding=True)
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    for name, o, n in BLOCKS:
        a, b = g[o:o + n], g_ref[o:o + n]
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b) + 1e-9, name


def test_mlp_training_reduces_loss(engine):
    from tensor_train_ring_decompositions_hpc_gpu_amd import learned as Lr
    X, T, info = Lr.make_dataset(8192, seed=11, workers=1)
    engine.mlp_init(5)
    Xd, Td = _dev(X), _dev(T)
    _, l0 = engine.mlp_gradient(Xd, Td)
    for it in range(300):
        i = (it * 512) % 8192
        engine.mlp_train_step(Xd[i:i + 512], Td[i:i + 512], 2e-3)
    _, l1 = engine.mlp_gradient(Xd, Td)
    assert l1 < 0.2 * l0, (l0, l1)


def test_features_kernel_matches_oracle(engine):
    import torch
    rng = np.random.default_rng(8)
    piv = 10.0 ** rng.uniform(-14, 3, 64)
    piv[50:] = 0.0
    x = engine.mlp_features(torch.from_numpy(piv).cuda(), 3e-7, 1536, 110592).cpu().numpy()
    ref = om.features(piv, 3e-7, 1536, 110592)
    assert np.max(np.abs(x - ref)) < 1e-6


def test_fused_predict_matches_features_and_forward(engine):
    """The one-launch prediction a step makes equals features + forward bit for bit, and the oracle's
    FP64 network with binary16 operands within FP32 accumulation error."""
    import torch
    rng = np.random.default_rng(9)
    engine.mlp_init(3)
    for trial in range(4):
        piv = 10.0 ** rng.uniform(-14, 3, 64)
        piv[40 + 5 * trial:] = 0.0
        dp = torch.from_numpy(piv).cuda()
        x = engine.mlp_features(dp, 1e-6 * (trial + 1), 1024, 1 << 20)
        y_ref = engine.mlp_forward(x[None, :]).cpu().numpy()[0]
        y = engine.mlp_predict(dp, 1e-6 * (trial + 1), 1024, 1 << 20).cpu().numpy()
        assert np.array_equal(y, y_ref), (y, y_ref)
        yo, _ = om.forward(engine.mlp_get_params(), om.features(piv, 1e-6 * (trial + 1), 1024, 1 << 20)[None, :])
        assert np.max(np.abs(y - yo[0])) < 1e-4 * max(1.0, np.max(np.abs(yo)))


@pytest.mark.parametrize("m,n,rank", [(96, 5000, 20), (300, 20000, 64), (128, 200000, 8)])
def test_probe_matches_oracle(engine, m, n, rank):
    rng = np.random.default_rng(m)
    U = np.linalg.qr(rng.standard_normal((m, rank)))[0]
    V = np.linalg.qr(rng.standard_normal((n, rank)))[0]
    s = np.arange(1, rank + 1) ** -1.5
    A = ((U * s) @ V.T).astype(np.float32)
    dA = _dev(A.reshape(-1, order="F"))
    piv = engine.probe(dA, m, n, seed=1234, stream=3)
    ref = om.probe_pivots(A.astype(np.float64), 1234, 3)
    tot = ref.sum()
    big = ref > 1e-6 * tot
    assert np.allclose(piv[big], ref[big], rtol=1e-4)
    x = om.features(piv, 1e-6, m, n)
    xr = om.features(ref, 1e-6, m, n)
    assert np.max(np.abs(x - xr)[:32][xr[:32] > -0.6]) < 0.02


# ------------------------------------------------------------------------------------------------
# the trained network in the sweep (learned-rank workload, smaller tensors)
# ------------------------------------------------------------------------------------------------
def _spectral(seed, dims, rmax=48):
    from tensor_train_ring_decompositions_hpc_gpu_amd import learned as Lr
    alphas = np.random.default_rng(seed).uniform(0.5, 3.0, len(dims) - 1)
    A, r = Lr.spectral_tt(dims, rmax, alphas, seed)
    return A, r


def test_learned_ranks_match_exact_truncation(engine):
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    if not os.path.exists(L.MODEL_PATH):
        pytest.skip("no trained rank model")
    engine.load_rank_model()
    dims, eps = [16, 16, 16, 16, 16], 1e-3
    for seed in (1, 2):
        A, r = _spectral(seed, dims)
        ceil = [1] + [min(48, x) for x in r[1:-1]] + [1]
        cores, sig, rep = engine.tt_svd(_dev(A), dims, ceil, s=8, q=2, precision="fp32", rank_mode="learned",
                                        eps=eps)
        _, _, _, used = ttsvd.tt_svd(A.astype(np.float64), dims, ceil, method="exact", rank_mode="delta", eps=eps)
        assert all(abs(a - b) <= 1 for a, b in zip(rep["ranks"], used)), (rep["ranks"], used)
        err = engine.rel_error(cores, dims, rep["ranks"], _dev(A))
        assert err <= 1.1 * eps, err


def test_learned_gain_over_adaptive_ladder(engine):
    """The trained network picks each randomized step's sketch width from the probe; without it the
    adaptive range finder climbs 16, 32, ... until its check passes.  Both meet eps; the network
    moves fewer bytes (probe included) over a set of tensors."""
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    if not os.path.exists(L.MODEL_PATH):
        pytest.skip("no trained rank model")
    engine.load_rank_model()
    dims, eps = [32, 32, 32, 32, 32], 1e-3
    tot = {"delta": [0, 0, 0.0], "learned": [0, 0, 0.0]}
    fallbacks = 0
    for seed in range(6):
        A, r = _spectral(seed, dims)
        dA = _dev(A)
        ceil = [1] + [min(48, x) for x in r[1:-1]] + [1]
        for mode in ("delta", "learned"):
            cores, sig, rep = engine.tt_svd(dA, dims, ceil, s=8, q=2, precision="bf16", rank_mode=mode, eps=eps,
                                            flags=L.FLAG_TIMING)
            err = engine.rel_error(cores, dims, rep["ranks"], dA)
            assert err <= 1.2 * eps, (seed, mode, err)
            tot[mode][0] += rep["bytes_alg"]
            tot[mode][1] += sum(rep["passes"])
            tot[mode][2] += rep["total_ms"]
            if mode == "learned":
                fallbacks += sum(rep["fallback"])
    print("adaptive only: %d bytes, %d passes, %.2f ms; learned: %d bytes, %d passes, %.2f ms; fallbacks %d"
          % (*tot["delta"], *tot["learned"], fallbacks))
    assert tot["learned"][0] < tot["delta"][0]
