"""Pins the FP64 CPU oracle (oracle/) against published known answers, numpy/LAPACK routines and
closed forms.  CPU only."""
import numpy as np
import pytest

from oracle import philox, ttsvd

# ------------------------------------------------------------------------------------------------
# Philox4x32-10: the Random123 known-answer vectors (Salmon et al., SC'11, kat_vectors)
# ------------------------------------------------------------------------------------------------
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    c = tuple(np.array([x], dtype=np.uint32) for x in ctr)
    k = tuple(np.array([x], dtype=np.uint32) for x in key)
    out = philox.philox4x32_10(c, k)
    assert tuple(int(o[0]) for o in out) == expect


def test_log_matches_libm():
    u = np.concatenate([np.linspace(2.0 ** -24, 1.0, 200001), 2.0 ** -np.arange(1, 25)]).astype(np.float32)
    err = np.abs(philox.log_f32(u).astype(np.float64) - np.log(u.astype(np.float64)))
    assert err.max() < 2e-6


def test_sincos_matches_libm():
    xb = (np.arange(0, 1 << 24, 97, dtype=np.uint64) << np.uint64(8)).astype(np.uint32)
    c, s = philox.sincos_quadrant(xb)
    ang = 2 * np.pi * (xb >> 8).astype(np.float64) / 2.0 ** 24
    assert np.abs(c - np.cos(ang)).max() < 2e-6
    assert np.abs(s - np.sin(ang)).max() < 2e-6


def test_normals_moments():
    z = philox.normal_flat(1134, 7, 1 << 20).astype(np.float64)
    assert abs(z.mean()) < 5e-3
    assert abs(z.var() - 1.0) < 5e-3
    # fourth moment of a Gaussian is 3
    assert abs(np.mean(z ** 4) - 3.0) < 5e-2
    # element e comes from normal4(seed, stream, e // 4, 0)[e % 4]
    z4 = philox.normal4(1134, 7, np.arange(3, dtype=np.uint64), np.zeros(3, dtype=np.uint32)).reshape(-1)
    assert np.array_equal(z4.astype(np.float32), philox.normal_flat(1134, 7, 12))


def test_round_bf16_is_round_to_nearest_even():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(100000).astype(np.float32),
                        # exact ties: low 16 bits 0x8000, both parities of bit 16
                        (np.arange(1000, dtype=np.uint32) * 0x10000 + 0x3F808000).view(np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(philox.round_bf16(x), ref)


def test_omega_entries_exact_in_low_precision():
    om = philox.omega(1234, 3, 4096, 42)
    assert om.shape == (4096, 42)
    assert np.array_equal(om.astype(np.float16).astype(np.float32), om)
    assert np.array_equal(philox.round_bf16(om), om)
    assert np.all((om == 0) | (np.abs(om) >= 2.0 ** -17))
    # columns beyond ell are not part of the sketch; the first ell columns do not depend on ell
    assert np.array_equal(philox.omega(1234, 3, 4096, 40), om[:, :40])


# ------------------------------------------------------------------------------------------------
# dense pieces against LAPACK
# ------------------------------------------------------------------------------------------------
def test_gram_eig_step_matches_svd():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((12, 300)) * np.logspace(0, -3, 12)[:, None]
    G, sig, An = ttsvd.gram_eig_step(A, 5)
    s_ref = np.linalg.svd(A, compute_uv=False)
    assert np.allclose(sig[: len(s_ref)], s_ref, rtol=1e-9, atol=1e-12)
    U = np.linalg.svd(A)[0][:, :5]
    assert np.allclose(G @ G.T, U @ U.T, atol=1e-9)
    assert np.allclose(An, G.T @ A)
    # gauge: the largest |entry| of every column is positive
    idx = np.argmax(np.abs(G), axis=0)
    assert np.all(G[idx, np.arange(5)] > 0)


def test_gram_eig_step_zero_pads_when_rank_exceeds_rows():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((6, 50))
    G, sig, An = ttsvd.gram_eig_step(A, 9)
    assert G.shape == (6, 9) and An.shape == (9, 50)
    assert np.allclose(G[:, 6:], 0) and np.allclose(An[6:], 0)
    assert np.allclose(G @ An, A, atol=1e-12)


def test_cholqr2_orthogonal_and_consistent():
    rng = np.random.default_rng(3)
    Y = rng.standard_normal((500, 20)) @ np.diag(np.logspace(0, -6, 20))
    Q, R = ttsvd.cholqr2(Y)
    assert np.abs(Q.T @ Q - np.eye(20)).max() < 1e-13
    assert np.allclose(Q @ R, Y, atol=1e-12)
    assert np.allclose(np.tril(R, -1), 0) and np.all(np.diag(R) > 0)
    Qh, Rh = ttsvd.qr_pos(Y)
    assert np.allclose(Q, Qh, atol=1e-8)


def test_rsvd_exact_when_rank_fits_sketch():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((80, 6)) @ rng.standard_normal((6, 500))
    U, sig, SVT = ttsvd.rsvd(A, 6, 4, 0, seed=1234)
    assert np.allclose(U @ SVT, A, atol=1e-10)
    assert np.allclose(sig[:6], np.linalg.svd(A, compute_uv=False)[:6], rtol=1e-10)
    assert np.abs(U.T @ U - np.eye(10)).max() < 1e-12


def test_rsvd_power_iterations_reduce_error():
    # a planted flat tail: without power iterations the randomized range misses more of the top-r
    # subspace than with them
    rng = np.random.default_rng(5)
    m, n, r = 200, 3000, 10
    U = np.linalg.qr(rng.standard_normal((m, m)))[0]
    V = np.linalg.qr(rng.standard_normal((n, m)))[0]
    s = np.concatenate([np.logspace(0, -1, r), np.full(m - r, 3e-2)])
    A = (U * s) @ V.T
    def err(q):
        Uq, sq, SVTq = ttsvd.rsvd(A, r, 5, q, seed=99)
        return np.linalg.norm(A - Uq[:, :r] @ SVTq[:r]) / np.linalg.norm(A)
    best = np.sqrt(np.sum(s[r:] ** 2)) / np.linalg.norm(s)
    assert err(2) < err(0)
    assert err(2) < 1.01 * best


def test_delta_rank_hand_computed():
    lam = np.array([100.0, 10.0, 1.0, 0.1, 0.01])
    total = lam.sum()
    assert ttsvd.delta_rank(lam, total, 0.2, 5) == 3      # tail(3) = 0.11 <= 0.2
    assert ttsvd.delta_rank(lam, total, 0.1, 5) == 4      # tail(3) = 0.11 > 0.1, tail(4) = 0.01
    assert ttsvd.delta_rank(lam, total, 0.0, 5) == 5
    assert ttsvd.delta_rank(lam, total, 1e9, 5) == 1
    assert ttsvd.delta_rank(lam, total, 0.0, 2) == 2      # cap
    # unknown tail beyond the captured values counts through total2
    assert ttsvd.delta_rank(lam, total + 5.0, 1.0, 5) == 5


# ------------------------------------------------------------------------------------------------
# permutations: brute force of the reference formulas
# ------------------------------------------------------------------------------------------------
def test_permute_231_bruteforce():
    r0, r1, n = 3, 4, 5
    T1 = np.arange(r0 * r1 * n, dtype=np.float64)
    T2 = ttsvd.permute_231(T1, r0, r1, n)
    ref = np.empty_like(T1)
    for a in range(r0):
        for b in range(r1):
            for j in range(n):
                # tensorToMode231 (trdec6_singleGPU.cu:49-58)
                ref[b + r1 * j + r1 * n * a] = T1[a + r0 * b + r0 * r1 * j]
    assert np.array_equal(T2, ref)
    assert np.array_equal(ttsvd.permute_312(T2, r1, n, r0), T1)


def test_permute_312_bruteforce():
    r1, n, r0 = 2, 3, 4
    T1 = np.arange(r0 * r1 * n, dtype=np.float64)
    T2 = ttsvd.permute_312(T1, r1, n, r0)
    ref = np.empty_like(T1)
    for i in range(r0 * r1 * n):
        tube, rem = divmod(i, r1 * n)
        row, col = rem % r1, rem // r1
        # tensorToMode3_host (trdec6_singleGPU.cu:166-173): m = r1, n = n, k = r0
        ref[r0 * col * r1 + row * r0 + tube] = T1[tube * r1 * n + col * r1 + row]
    assert np.array_equal(T2, ref)


# ------------------------------------------------------------------------------------------------
# sweeps: exact-rank inputs reconstruct to roundoff; contractions against einsum
# ------------------------------------------------------------------------------------------------
def test_tt_full_matches_einsum():
    dims, ranks = [3, 4, 5], [1, 2, 3, 1]
    cores = ttsvd.make_tt_cores(dims, ranks)
    G = [c.astype(np.float64).reshape(ranks[k], dims[k], ranks[k + 1], order="F") for k, c in enumerate(cores)]
    ref = np.einsum("aib,bjc,ckd->ijk", *G).reshape(-1, order="F")
    assert np.allclose(ttsvd.tt_full(cores, dims, ranks), ref, atol=1e-12)


def test_tr_full_matches_einsum():
    dims, ranks = [4, 5, 6], [2, 2, 3, 2]
    cores = ttsvd.make_tr_cores(dims, ranks)
    G1 = cores[0].astype(np.float64).reshape(4, 2, 2, order="F")    # [i, a, b], column a + r0 b
    G2 = cores[1].astype(np.float64).reshape(2, 5, 3, order="F")    # [b, j, c]
    G3 = cores[2].astype(np.float64).reshape(3, 6, 2, order="F")    # [c, k, a]
    ref = np.einsum("iab,bjc,cka->ijk", G1, G2, G3).reshape(-1, order="F")
    assert np.allclose(ttsvd.tr_full(cores, dims, ranks), ref, atol=1e-12)


@pytest.mark.parametrize("method", ["exact", "rsvd"])
def test_tt_svd_exact_rank_reconstructs(method):
    dims, ranks = [6, 7, 8, 5], [1, 3, 4, 2, 1]
    A = ttsvd.tt_full(ttsvd.make_tt_cores(dims, ranks), dims, ranks)
    cores, sig, paths, used = ttsvd.tt_svd(A, dims, ranks, s=3, q=1, method=method)
    assert used == ranks
    assert ttsvd.rel_error(A, ttsvd.tt_full(cores, dims, used)) < 1e-12
    for k in range(len(dims) - 1):          # left-orthogonal cores
        assert np.abs(cores[k].T @ cores[k] - np.eye(ranks[k + 1])).max() < 1e-12


def test_tt_svd_truncation_error_bound():
    rng = np.random.default_rng(6)
    dims = [5, 6, 7, 4]
    A = rng.standard_normal(np.prod(dims))
    ranks = [1, 3, 4, 3, 1]
    cores, sig, paths, used = ttsvd.tt_svd(A, dims, ranks, method="exact")
    err = ttsvd.rel_error(A, ttsvd.tt_full(cores, dims, used))
    # Oseledets 2011, Thm 2.2: ||A - TT||^2 <= sum_k tail_k^2
    bound = np.sqrt(sum(np.sum(sig[k][ranks[k + 1]:] ** 2) for k in range(3))) / np.linalg.norm(A)
    assert err <= bound * (1 + 1e-9)
    # step 1 singular values are those of the first unfolding
    s1 = np.linalg.svd(A.reshape(dims[0], -1, order="F"), compute_uv=False)
    assert np.allclose(sig[0][: len(s1)], s1, rtol=1e-10)


def test_tt_svd_delta_ranks():
    dims, ranks = [6, 7, 8, 5], [1, 3, 4, 2, 1]
    A = ttsvd.tt_full(ttsvd.make_tt_cores(dims, ranks), dims, ranks)
    cores, sig, paths, used = ttsvd.tt_svd(A, dims, [1, 30, 30, 30, 1], method="exact", rank_mode="delta",
                                            eps=1e-6)
    assert used == ranks
    assert ttsvd.rel_error(A, ttsvd.tt_full(cores, dims, used)) < 1e-6


def test_tr_svd_exact_ranks_reconstruct():
    dims, true = [4, 5, 6], [2, 2, 3, 2]
    T = ttsvd.tr_full(ttsvd.make_tr_cores(dims, true), dims, true)
    cores, sig, paths, used = ttsvd.tr_svd(T, dims, [2, 2, 10, 2])
    assert [c.shape for c in cores] == [(4, 4), (10, 10), (10, 12)]
    assert ttsvd.rel_error(T, ttsvd.tr_full(cores, dims, used)) < 1e-12


def test_tr_svd_order6_exact_at_r0_squared_ranks():
    # the first step keeps an arbitrary basis of the r0 r1 column space, so the chain that follows
    # carries the ring index twice: inner ranks r0^2 r_k reproduce a ring of ranks r_k exactly,
    # smaller ones do not
    dims, true = [8] * 6, [2] * 7
    T = ttsvd.tr_full(ttsvd.make_tr_cores(dims, true, seed=1134), dims, true)
    cores, _, _, used = ttsvd.tr_svd(T, dims, [2, 2, 8, 8, 8, 8, 2])
    assert ttsvd.rel_error(T, ttsvd.tr_full(cores, dims, used)) < 1e-12
    cores, _, _, used = ttsvd.tr_svd(T, dims, [2, 2, 4, 4, 4, 4, 2])
    assert ttsvd.rel_error(T, ttsvd.tr_full(cores, dims, used)) > 1e-3


def test_tr_svd_delta_ranks():
    """delta-ranked TR-SVD: the first step keeps r0 * ceil(R / r0) columns for its delta rank R (a
    first unfolding of known rank R = r0_true * r1_true), every later step the delta rank; the error
    stays within eps; on exact steps (no range finder, whose checked widths differ from the fixed ones)
    a tiny eps with the fixed ranks as ceilings reproduces the fixed sweep."""
    dims, true = [8, 9, 10, 11], [2, 3, 4, 3, 2]
    T = ttsvd.tr_full(ttsvd.make_tr_cores(dims, true, seed=5), dims, true)
    T = ttsvd.add_noise(T, 1e-6, seed=6)
    # the first unfolding has rank 6 (up to the noise): r0 = 2 -> r1 = 3, r0 = 4 -> r1 = 2 (8 columns)
    for r0, r1 in ((2, 3), (4, 2)):
        cores, _, _, used = ttsvd.tr_svd(T, dims, [r0, 8, 48, 48, r0], s=6, q=1, rank_mode="delta", eps=1e-3)
        assert used[0] == used[4] == r0 and used[1] == r1, used
        assert cores[0].shape == (8, r0 * r1)
        assert ttsvd.rel_error(T, ttsvd.tr_full(cores, dims, used)) <= 1e-3
    fixed = [2, 3, 5, 4, 2]
    c_f, s_f, _, u_f = ttsvd.tr_svd(T, dims, fixed, method="exact")
    c_d, s_d, _, u_d = ttsvd.tr_svd(T, dims, fixed, method="exact", rank_mode="delta", eps=1e-12)
    assert u_d == u_f == fixed
    for a, b in zip(c_f, c_d):
        assert np.array_equal(a, b)


def test_noise_and_generator_definitions():
    dims, ranks = [5, 6, 7], [1, 2, 3, 1]
    cores = ttsvd.make_tt_cores(dims, ranks, std=0.5)
    assert np.array_equal(cores[1].reshape(-1, order="F"), philox.normal_flat(ttsvd.DATA_SEED, 0x101, 36) * np.float32(0.5))
    A = ttsvd.tt_full(cores, dims, ranks).astype(np.float32)
    An = ttsvd.add_noise(A, 1e-3)
    rel = np.linalg.norm(An.astype(np.float64) - A) / np.linalg.norm(A)
    assert 0.8e-3 < rel < 1.2e-3


# ------------------------------------------------------------------------------------------------
# adaptive range finder (a-posteriori check on held-out Gaussian test columns)
# ------------------------------------------------------------------------------------------------
def _spectral(m, n, sig, seed):
    rng = np.random.default_rng(seed)
    U = np.linalg.qr(rng.standard_normal((m, len(sig))))[0]
    V = np.linalg.qr(rng.standard_normal((n, len(sig))))[0]
    return (U * np.asarray(sig)) @ V.T


def test_range_resid2_is_unbiased():
    # E ||(I - QQ^T) A w||^2 = ||(I - QQ^T) A||_F^2 for w ~ N(0, I): average over many test columns
    rng = np.random.default_rng(3)
    A = _spectral(60, 400, np.linspace(1.0, 0.05, 30), 4)
    Q = np.linalg.qr(rng.standard_normal((60, 12)))[0]
    W = rng.standard_normal((400, 20000))
    est = ttsvd.range_resid2(Q, A @ W)
    R = A - Q @ (Q.T @ A)
    exact = np.sum(R * R) / np.sum(A * A)
    assert abs(est - exact) < 0.03 * exact


def test_range_resid2_closed_forms():
    rng = np.random.default_rng(5)
    Q = np.linalg.qr(rng.standard_normal((40, 6)))[0]
    Yin = Q @ rng.standard_normal((6, 8))          # inside span(Q): zero residual
    assert ttsvd.range_resid2(Q, Yin) < 1e-28
    P = np.linalg.qr(np.hstack([Q, rng.standard_normal((40, 8))]))[0][:, 6:]
    Yout = P @ rng.standard_normal((8, 8))         # orthogonal to span(Q): residual ratio 1
    assert abs(ttsvd.range_resid2(Q, Yout) - 1.0) < 1e-12


def test_adaptive_widths_ladder():
    assert ttsvd.adaptive_widths(42, "fixed") == [42]
    assert ttsvd.adaptive_widths(42, "delta") == [16, 32, 42]
    assert ttsvd.adaptive_widths(56, "delta") == [16, 32, 56]
    assert ttsvd.adaptive_widths(12, "delta") == [12]
    assert ttsvd.adaptive_widths(42, "delta", pred_ell=20) == [20, 32, 42]
    assert ttsvd.adaptive_widths(42, "fixed", pred_ell=40) == [40, 42]


def test_adaptive_range_exact_rank_needs_no_power_iteration():
    A = _spectral(80, 500, np.logspace(0, -3, 10), 7)
    Q, w, it, rel2, checks = ttsvd.adaptive_range(A, [16], 8, 2, 1e-12, 1234, 1)
    assert (w, it, checks) == (16, 0, 1) and rel2 < 1e-26
    assert np.allclose(Q @ (Q.T @ A), A, atol=1e-12)


def test_adaptive_range_grows_to_the_rank():
    # rank 24: a 16-column basis misses a third of the spectrum, 32 columns capture it exactly
    A = _spectral(80, 600, np.ones(24), 8)
    Q, w, it, rel2, checks = ttsvd.adaptive_range(A, [16, 32, 42], 8, 2, 1e-10, 1234, 2)
    assert (w, it, checks) == (32, 0, 2) and rel2 < 1e-24


def test_adaptive_range_power_iteration_then_stop():
    # 8 strong directions over a flat tail: q = 0 leaves a tail-sized residual that a power
    # iteration removes, and the check stops after the first one
    sig = np.concatenate([np.ones(8), np.full(200, 3e-3)])
    A = _spectral(300, 800, sig, 9)
    Q0, _, _, r0, _ = ttsvd.adaptive_range(A, [10], 8, 0, 0.0, 1234, 3)
    Q1, _, _, r1, _ = ttsvd.adaptive_range(A, [10], 8, 1, 0.0, 1234, 3)
    assert r1 < 0.3 * r0
    best = 198 * 9e-6 / (8 + 200 * 9e-6)          # the optimal 10-dimensional range leaves 198 tail values
    assert abs(r1 - best) < 0.05 * best
    tol2 = np.sqrt(r0 * r1)
    Q, w, it, rel2, checks = ttsvd.adaptive_range(A, [10], 8, 3, tol2, 1234, 3)
    assert (it, checks) == (1, 2) and rel2 == r1


def test_adaptive_tt_svd_matches_fixed_width_sweep():
    # an exact-rank tensor: the adaptive sweep stops at q = 0 and equals the q = 0 sweep
    dims, ranks = [6, 8, 8, 8, 8], [1, 5, 6, 6, 5, 1]
    cores = ttsvd.make_tt_cores(dims, ranks, std=1.0, seed=21)
    A = ttsvd.tt_full(cores, dims, ranks)
    infos = []
    c1, s1, p1, u1 = ttsvd.tt_svd(A, dims, ranks, s=6, q=2, adaptive=True, infos=infos)
    c0, s0, p0, u0 = ttsvd.tt_svd(A, dims, ranks, s=6, q=0)
    assert u1 == u0 and p1 == p0
    for k, inf in enumerate(infos):
        if p1[k] == "rsvd":
            assert inf["power_iters"] == 0 and inf["checks"] == 1
        assert np.allclose(s1[k][:ranks[k + 1]], s0[k][:ranks[k + 1]], rtol=1e-10)
