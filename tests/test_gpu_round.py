"""TT rounding on the GPU (ttx_tt_round) against the FP64 oracle's (oracle/ttsvd.py tt_round): ranks
identical, bond singular values within 1e-5 of the largest, reconstruction errors within the budget;
the sum of a TT with itself rounds back to its ranks; a sweep's over-ranked cores round down to the
data's ranks."""
import numpy as np
import pytest

from oracle import ttsvd

pytestmark = pytest.mark.gpu


def _dev(cores):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(np.asarray(c, dtype=np.float32).reshape(-1, order="F"))).cuda()
            for c in cores]


def _graded(dims, ranks, seed, rate=0.3):
    rng = np.random.default_rng(seed)
    c = [rng.standard_normal((ranks[k] * dims[k], ranks[k + 1])) for k in range(len(dims))]
    for k in range(len(dims) - 1):
        c[k] = c[k] * (rate ** np.arange(ranks[k + 1]))[None, :]
    return [x.astype(np.float32) for x in c]


def test_round_sum_of_tt_with_itself(engine):
    dims, ranks = [12, 10, 14, 9], [1, 5, 7, 4, 1]
    A, cores = engine.generate(dims, ranks, std=1.0, noise_rel=0.0, seed=41, return_cores=True)
    host = [c.cpu().numpy().astype(np.float64) for c in cores]
    cs, rs = ttsvd.tt_add(host, host, dims, ranks, ranks)
    # the sum's cores are exactly rank deficient: the directions beyond the summands' ranks come out at
    # the FP32 rounding level of the orthogonalized cores (FP64 Grams), far below eps = 1e-5
    out, rn, sig = engine.tt_round(_dev(cs), dims, rs, 1e-5)
    print("round sum ranks", rs, "->", rn, "bond 1 sigma", sig[0][:8], flush=True)
    assert rn == ranks
    assert sig[0][5] < 1e-6 * sig[0][0]
    full = engine.full(out, dims, rn).cpu().numpy().astype(np.float64)
    ref = 2.0 * A.cpu().numpy().astype(np.float64)
    assert ttsvd.rel_error(ref, full) < 1e-5


@pytest.mark.parametrize("eps", [1e-2, 1e-4])
def test_round_matches_oracle(engine, eps):
    dims, ranks = [10, 12, 9, 11], [1, 8, 12, 6, 1]
    c = _graded(dims, ranks, 7)
    out, rn, sig = engine.tt_round(_dev(c), dims, ranks, eps)
    cr, rr, sr = ttsvd.tt_round(c, dims, ranks, eps)
    print("round eps", eps, "ranks", rn, "oracle", rr, flush=True)
    assert rn == rr
    for k in range(len(dims) - 1):
        n = min(ranks[k + 1], len(sr[k]))
        assert np.allclose(sig[k][:n], sr[k][:n], rtol=0, atol=1e-5 * sr[k][0]), k
    A = ttsvd.tt_full(c, dims, ranks)
    err = ttsvd.rel_error(A, engine.full(out, dims, rn).cpu().numpy())
    err_ref = ttsvd.rel_error(A, ttsvd.tt_full(cr, dims, rr))
    assert err <= eps * 1.001 and abs(err - err_ref) <= 1e-5 + 1e-2 * err_ref, (err, err_ref)


def test_round_rank_cap(engine):
    dims, ranks = [10, 12, 9, 11], [1, 8, 12, 6, 1]
    c = _graded(dims, ranks, 8)
    cap = [1, 3, 4, 2, 1]
    out, rn, _ = engine.tt_round(_dev(c), dims, ranks, 1e-9, max_ranks=cap)
    cr, rr, _ = ttsvd.tt_round(c, dims, ranks, 1e-9, max_ranks=cap)
    assert rn == rr == cap
    A = ttsvd.tt_full(c, dims, ranks)
    err = ttsvd.rel_error(A, engine.full(out, dims, rn).cpu().numpy())
    err_ref = ttsvd.rel_error(A, ttsvd.tt_full(cr, dims, rr))
    assert abs(err - err_ref) <= 1e-5 + 1e-3 * err_ref, (err, err_ref)


def test_round_sweep_cores_down_to_data_ranks(engine):
    """A TT-SVD at ranks above the data's keeps noise directions; rounding its cores with eps above the
    noise drops them: the ranks become the data's and the error stays within eps."""
    dims, true_r = [16, 16, 16, 16], [1, 4, 6, 3, 1]
    A = engine.generate(dims, true_r, std=1.0, noise_rel=1e-7, seed=43)
    big = [1, 12, 20, 10, 1]
    cores, _, rep = engine.tt_svd(A, dims, big, s=10, q=2, precision="fp32")
    out, rn, _ = engine.tt_round(cores, dims, rep["ranks"], 1e-4)
    assert rn == true_r
    assert engine.rel_error(out, dims, rn, A) < 1e-4
