"""The oracle's TT rounding (oracle/ttsvd.py tt_round, tt_add), pinned by what the mathematics fixes:
an exact TT sum rounds back to the summands' ranks and reconstructs the sum; with no truncation the
bond singular values are those of the full tensor's unfoldings; truncated rounding gives the ranks and
error of the exact TT-SVD of the full tensor with the same delta rule."""
import numpy as np

from oracle import ttsvd


def _tt(dims, ranks, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((ranks[k] * dims[k], ranks[k + 1])) for k in range(len(dims))]


def _unfold_svals(A, dims, k):
    m = int(np.prod(dims[:k + 1]))
    return np.linalg.svd(A.reshape(m, -1, order="F"), compute_uv=False)


def test_sum_of_a_tt_with_itself_rounds_to_its_ranks():
    dims, ranks = [5, 6, 7, 4], [1, 3, 4, 2, 1]
    c = _tt(dims, ranks, 1)
    A = ttsvd.tt_full(c, dims, ranks)
    cs, rs = ttsvd.tt_add(c, c, dims, ranks, ranks)
    assert rs == [1, 6, 8, 4, 1]
    assert np.allclose(ttsvd.tt_full(cs, dims, rs), 2 * A, rtol=0, atol=1e-12 * np.abs(A).max())
    cr, rr, _ = ttsvd.tt_round(cs, dims, rs, 1e-10)
    assert rr == ranks
    assert ttsvd.rel_error(2 * A, ttsvd.tt_full(cr, dims, rr)) < 1e-12


def test_bond_singular_values_are_the_unfoldings():
    dims, ranks = [4, 5, 6, 3], [1, 4, 5, 3, 1]
    c = _tt(dims, ranks, 2)
    A = ttsvd.tt_full(c, dims, ranks)
    _, rr, sig = ttsvd.tt_round(c, dims, ranks, 0.0)
    assert rr == ranks
    for k in range(len(dims) - 1):
        ref = _unfold_svals(A, dims, k)[:ranks[k + 1]]
        assert np.allclose(sig[k][:ranks[k + 1]], ref, rtol=1e-10, atol=1e-12 * ref[0]), k


def test_truncated_rounding_matches_exact_tt_svd_of_the_full_tensor():
    dims, ranks = [6, 7, 5, 6], [1, 5, 8, 4, 1]
    c = _tt(dims, ranks, 3)
    # graded bonds: scale core columns so each unfolding's spectrum decays
    for k in range(len(dims) - 1):
        c[k] = c[k] * (0.3 ** np.arange(ranks[k + 1]))[None, :]
    A = ttsvd.tt_full(c, dims, ranks)
    for eps in (1e-2, 1e-4):
        cr, rr, _ = ttsvd.tt_round(c, dims, ranks, eps)
        err = ttsvd.rel_error(A, ttsvd.tt_full(cr, dims, rr))
        assert err <= eps * (1 + 1e-9)
        _, _, _, used = ttsvd.tt_svd(A, dims, ranks, method=ttsvd.EXACT, rank_mode="delta", eps=eps)
        assert rr == list(used), (eps, rr, used)
    # a rank cap wins over eps
    cr, rr, _ = ttsvd.tt_round(c, dims, ranks, 1e-12, max_ranks=[1, 2, 2, 2, 1])
    assert rr == [1, 2, 2, 2, 1]
