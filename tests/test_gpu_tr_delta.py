"""delta-ranked TR-SVD against the FP64 oracle (oracle/ttsvd.py tr_svd, rank_mode="delta").

Every step truncates by the delta rule with delta^2 = eps^2 / (d-1) ||A||_F^2; the ring rank r0 stays as
given and the first step keeps r0 * ceil(R / r0) columns for its delta rank R.  Checked: the ranks the
engine reports equal the oracle's, the relative error stays within eps and within the arithmetic's
tolerance of the oracle's, and the kept singular values agree.  Cases: an exact first step (with and
without rows beyond its rank bound), and a randomized first step (its remainder permuted to the 231
layout afterwards)."""
import numpy as np
import pytest

from oracle import ttsvd

pytestmark = pytest.mark.gpu

TOL = {"fp32": dict(sig=1e-5, err=1e-5), "bf16": dict(sig=1e-4, err=1e-5)}


def host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("dims,true,ceil,first_path", [
    ([8, 9, 10, 11], [2, 3, 4, 3, 2], [2, 8, 48, 48, 2], 0),    # exact first step, r1 = 3
    ([8, 9, 10, 11], [2, 3, 4, 3, 2], [4, 8, 48, 48, 4], 0),    # exact first step, r0 = 4 -> r1 = 2
    ([80, 6, 7, 5], [2, 3, 4, 3, 2], [2, 8, 30, 30, 2], 2),     # randomized first step, permuted remainder
])
def test_tr_svd_delta_matches_oracle(engine, prec, dims, true, ceil, first_path):
    eps = 1e-3
    A = engine.generate(dims, true, std=1.0, noise_rel=1e-6, seed=71, tr=True)
    print("tr delta", dims, ceil, prec, flush=True)
    cores, sig, rep = engine.tr_svd(A, dims, ceil, s=6, q=1, precision=prec, rank_mode="delta", eps=eps)
    An = host(A).astype(np.float64)
    c_ref, s_ref, p_ref, used = ttsvd.tr_svd(An, dims, ceil, s=6, q=1, rank_mode="delta", eps=eps)
    assert rep["path"][0] == first_path
    assert rep["ranks"] == used, (rep["ranks"], used)
    err = engine.rel_error(cores, dims, rep["ranks"], A, tr=True)
    err_ref = ttsvd.rel_error(An, ttsvd.tr_full(c_ref, dims, used))
    tol = TOL[prec]
    assert err <= eps and abs(err - err_ref) <= tol["err"] + 2e-2 * err_ref, (err, err_ref)
    for k in range(len(dims) - 1):
        r = used[0] * used[1] if k == 0 else used[k + 1]
        r = min(r, len(s_ref[k]))
        got, ref = sig[k][:r].astype(np.float64), s_ref[k][:r]
        # rounding r up to a multiple of r0 can keep noise-level directions: a Gram-based sigma is exact
        # to eps sigma_1^2 in sigma^2 only, so those compare in that form
        big = ref >= 1e-3 * ref[0]
        assert np.allclose(got[big], ref[big], rtol=tol["sig"], atol=tol["sig"] * ref[0]), k
        assert np.allclose(got[~big] ** 2, ref[~big] ** 2, rtol=0, atol=tol["sig"] * ref[0] ** 2), k


def test_tr_svd_delta_tiny_eps_is_fixed(engine):
    """With the fixed ranks as ceilings and a tiny eps the delta sweep keeps every ceiling: on exact
    steps (no range finder, whose checked widths differ from the fixed ones) the same cores as the
    fixed-rank sweep, bit for bit (same kernels, same plans)."""
    dims, ranks = [8, 9, 10, 6], [2, 3, 5, 4, 2]
    A = engine.generate(dims, ranks, std=1.0, noise_rel=1e-6, seed=72, tr=True)
    c_f, s_f, r_f = engine.tr_svd(A, dims, ranks, precision="fp32", method="exact")
    c_d, s_d, r_d = engine.tr_svd(A, dims, ranks, precision="fp32", method="exact", rank_mode="delta", eps=1e-12)
    assert r_d["ranks"] == r_f["ranks"] == ranks
    for (a, _), (b, _) in zip(c_f, c_d):
        assert np.array_equal(host(a), host(b))
