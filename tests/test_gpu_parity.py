"""Parity of the HIP engine (through the C ABI) with the FP64 CPU oracle on the same seeded inputs
and the same Philox sketch.

Tolerances (relative):
* generators and permutations: bit-exact;
* FP32 arithmetic: singular values 1e-5, cores / remainders 1e-4;
* BF16 arithmetic (hi/lo split, FP32 accumulation): singular values 1e-4, cores / remainders 1e-3;
* reconstruction errors: |err_gpu - err_oracle| <= 1e-5 + 2e-2 * err_oracle in FP32 and in BF16 (the
  split products keep about 16 significant bits per operand; their error on these inputs is a few
  1e-6 of ||A||).
"""
import numpy as np
import pytest

from oracle import philox, ttsvd

pytestmark = pytest.mark.gpu

TOL = {"fp32": dict(sig=1e-5, vec=1e-4, err=1e-5), "bf16": dict(sig=1e-4, vec=1e-3, err=1e-5)}


def host(t):
    return t.detach().cpu().numpy()


def fmat(t, rows, cols):
    return host(t)[: rows * cols].reshape(rows, cols, order="F").astype(np.float64)


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1, order="F"))).cuda()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def lowrank(m, n, r, seed, decay=2.0, noise=1e-6):
    rng = np.random.default_rng(seed)
    U = np.linalg.qr(rng.standard_normal((m, r)))[0]
    V = np.linalg.qr(rng.standard_normal((n, r)))[0]
    s = np.logspace(0, -decay, r)
    A = (U * s) @ V.T
    A += noise * np.linalg.norm(A) / np.sqrt(m * n) * rng.standard_normal((m, n))
    return A.astype(np.float32)


# ------------------------------------------------------------------------------------------------
def test_fill_normal_bit_exact(engine):
    z = host(engine.fill_normal(10007, 1134, 0x101, 0.5))
    ref = philox.normal_flat(1134, 0x101, 10007) * np.float32(0.5)
    assert np.array_equal(z.view(np.uint32), ref.astype(np.float32).view(np.uint32))


def test_omega_bit_exact(engine):
    om = host(engine.omega(3001, 42, 1234, 3)).reshape(3001, 42)
    ref = philox.omega(1234, 3, 3001, 42)
    assert np.array_equal(om.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("r0,r1,n", [(4, 16, 1000), (3, 5, 7), (2, 4, 12345), (8, 3, 64)])
def test_permute_231_bit_exact(engine, r0, r1, n):
    import torch
    x = torch.randn(r0 * r1 * n, device="cuda")
    y = host(engine.permute_231(x, r0, r1, n))
    ref = ttsvd.permute_231(host(x), r0, r1, n)
    assert np.array_equal(y.view(np.uint32), ref.astype(np.float32).view(np.uint32))


def test_generate_tt_matches_oracle(engine):
    dims, ranks = [6, 7, 8, 5], [1, 3, 4, 2, 1]
    A, cores = engine.generate(dims, ranks, std=0.5, noise_rel=1e-3, seed=1134, return_cores=True)
    ref_cores = ttsvd.make_tt_cores(dims, ranks, std=0.5, seed=1134)
    for c, rc in zip(cores, ref_cores):
        assert np.array_equal(host(c), rc.reshape(-1, order="F"))
    clean = ttsvd.tt_full(ref_cores, dims, ranks).astype(np.float32)
    ref = ttsvd.add_noise(clean, 1e-3, seed=1134)
    assert rel(host(A), ref) < 1e-6


def test_generate_tr_matches_oracle(engine):
    dims, ranks = [6, 5, 4, 3], [2, 3, 2, 4, 2]
    A, cores = engine.generate(dims, ranks, std=0.5, noise_rel=0.0, seed=77, tr=True, return_cores=True)
    ref_cores = ttsvd.make_tr_cores(dims, ranks, std=0.5, seed=77)
    for c, rc in zip(cores, ref_cores):
        assert np.array_equal(host(c), rc.reshape(-1, order="F"))
    assert rel(host(A), ttsvd.tr_full(ref_cores, dims, ranks)) < 1e-6


# ------------------------------------------------------------------------------------------------
def _check_step(engine, A, r, s, q, prec, expect_path):
    m, n = A.shape
    core, nxt, sig, path = engine.step(dev(A), m, n, r, s=s, q=q, seed=1234, stream_id=5, precision=prec)
    assert path == expect_path
    G_ref, sig_ref, N_ref, _ = ttsvd.step(A.astype(np.float64), r, s, q, 1234, 5)
    tol = TOL[prec]
    k = min(r, len(sig_ref))
    assert np.allclose(sig[:k], sig_ref[:k], rtol=tol["sig"], atol=tol["sig"] * sig_ref[0]), (sig[:k], sig_ref[:k])
    G = fmat(core, m, r)
    assert rel(G @ G.T, G_ref @ G_ref.T) < tol["vec"]
    assert rel(G, G_ref) < tol["vec"], "core differs beyond the gauge"
    assert rel(fmat(nxt, r, n), N_ref) < tol["vec"]


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_step_exact_short(engine, prec):
    # m <= ell: exact Gram path (ttdec6_singleGPU.cu:510-536)
    _check_step(engine, lowrank(24, 3000, 12, 11), 8, 20, 2, prec, 0)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_step_exact_wide(engine, prec):
    # n <= ell: orthonormal basis of the columns, then the common tail
    _check_step(engine, lowrank(500, 20, 20, 12, decay=1.0), 12, 10, 2, prec, 1)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("m,n,r,s,q", [(1024, 4000, 16, 8, 1), (300, 2500, 10, 6, 2), (96, 777, 5, 3, 0),
                                       (1024, 3000, 32, 10, 2)])
def test_step_randomized(engine, prec, m, n, r, s, q):
    _check_step(engine, lowrank(m, n, r, m + n, decay=1.5), r, s, q, prec, 2)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("m,n,r,s,q", [(64, 140000, 16, 8, 0), (48, 131100, 10, 6, 1)])
def test_step_randomized_long(engine, prec, m, n, r, s, q):
    # n >= 2^17 takes the one-row-per-lane remainder kernel (vector stores for r % 4 == 0, scalar
    # stores otherwise)
    _check_step(engine, lowrank(m, n, r, 29, decay=1.5), r, s, q, prec, 2)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_rsvd_matches_oracle(engine, prec):
    A = lowrank(700, 3000, 24, 3, decay=2.0)
    U, S, SVT = engine.rsvd(dev(A), 700, 3000, 24, s=8, q=1, seed=1234, precision=prec)
    U_ref, s_ref, SVT_ref = ttsvd.rsvd(A.astype(np.float64), 24, 8, 1, 1234, 0)
    tol = TOL[prec]
    assert np.allclose(host(S)[:24], s_ref[:24], rtol=tol["sig"])
    Ug = fmat(U, 700, 32)
    assert rel(Ug[:, :24], U_ref[:, :24]) < tol["vec"]
    assert rel(fmat(SVT, 32, 3000)[:24], SVT_ref[:24]) < tol["vec"]


# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_tt_svd_matches_oracle(engine, prec):
    dims, true = [8, 9, 10, 11, 6], [1, 4, 6, 6, 5, 1]
    A = engine.generate(dims, true, std=1.0, noise_rel=1e-6, seed=1134)
    ranks = true
    cores, sig, rep = engine.tt_svd(A, dims, ranks, s=4, q=1, precision=prec)
    An = host(A).astype(np.float64)
    c_ref, s_ref, p_ref, used = ttsvd.tt_svd(An, dims, ranks, s=4, q=1)
    tol = TOL[prec]
    assert rep["ranks"] == used
    assert [pp == 2 for pp in rep["path"]] == [p == "rsvd" for p in p_ref]
    for k in range(len(dims) - 1):
        r = ranks[k + 1]
        assert np.allclose(sig[k][:r], s_ref[k][:r], rtol=tol["sig"], atol=tol["sig"] * s_ref[k][0]), k
    err = engine.rel_error(cores, dims, used, A)
    err_ref = ttsvd.rel_error(An, ttsvd.tt_full(c_ref, dims, used))
    assert abs(err - err_ref) <= tol["err"] + 2e-2 * err_ref, (err, err_ref)
    for k in range(len(dims) - 1):
        g, shape = cores[k]
        assert rel(fmat(g, *shape), c_ref[k]) < 10 * tol["vec"], k


# FP16 range finder (binary16 sketch operands with power-of-two scales, FP32 accumulation): an
# emulation of it in the oracle moves the singular values of these cases by at most 3.4e-5 of sigma_1
# and the reconstruction error by at most 3.4e-5 (the binary16 floor of a well-posed ring); the
# projection and remainder stay FP32-accurate
TOL["fp16"] = dict(sig=1e-3, vec=1e-2, err=5e-4)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_tr_svd_matches_oracle(engine, prec):
    dims, true = [8, 9, 10, 6], [2, 3, 5, 4, 2]
    A = engine.generate(dims, true, std=1.0, noise_rel=1e-6, seed=99, tr=True)
    ranks = [2, 3, 8, 6, 2]
    cores, sig, rep = engine.tr_svd(A, dims, ranks, s=4, q=1, precision=prec)
    assert rep["path"] == [0, 2, 2]
    An = host(A).astype(np.float64)
    c_ref, s_ref, p_ref, used = ttsvd.tr_svd(An, dims, ranks, s=4, q=1)
    tol = TOL[prec]
    for k in range(len(dims) - 1):
        r = (ranks[0] * ranks[1]) if k == 0 else ranks[k + 1]
        r = min(r, len(s_ref[k]))
        assert np.allclose(sig[k][:r], s_ref[k][:r], rtol=tol["sig"], atol=tol["sig"] * s_ref[k][0]), k
    err = engine.rel_error(cores, dims, ranks, A, tr=True)
    err_ref = ttsvd.rel_error(An, ttsvd.tr_full(c_ref, dims, ranks))
    assert abs(err - err_ref) <= tol["err"] + 2e-2 * err_ref, (err, err_ref)


def test_tt_rel_error_matches_oracle(engine):
    dims, ranks = [6, 7, 8, 5], [1, 3, 4, 2, 1]
    A, cores = engine.generate(dims, ranks, std=1.0, noise_rel=1e-2, seed=5, return_cores=True)
    err = engine.rel_error(cores, dims, ranks, A)
    ref_cores = [host(c) for c in cores]
    ref = ttsvd.rel_error(host(A), ttsvd.tt_full(ref_cores, dims, ranks))
    assert abs(err - ref) < 1e-6 + 1e-4 * ref


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("dims,ranks,s,q", [
    ([8, 8, 8, 8, 8], [1, 8, 8, 8, 8, 1], 6, 1),      # full-rank step 1 folded into a randomized step 2
    ([4, 4, 8, 8], [1, 4, 4, 4, 1], 12, 1),          # folded into an exact (short) step 2
    ([4, 4, 4], [1, 4, 4, 1], 4, 1),                 # folded into an exact (wide) last step
    ([32, 16, 16, 8], [1, 32, 16, 8, 1], 10, 2),     # cfg2's shape class at a small size
])
def test_tt_svd_full_rank_first_step(engine, prec, dims, ranks, s, q):
    A = engine.generate(dims, ranks, std=1.0, noise_rel=1e-6, seed=4321)
    cores, sig, rep = engine.tt_svd(A, dims, ranks, s=s, q=q, precision=prec)
    An = host(A).astype(np.float64)
    c_ref, s_ref, p_ref, used = ttsvd.tt_svd(An, dims, ranks, s=s, q=q)
    tol = TOL[prec]
    assert rep["ranks"] == used
    for k in range(len(dims) - 1):
        r = ranks[k + 1]
        assert np.allclose(sig[k][:r], s_ref[k][:r], rtol=tol["sig"], atol=tol["sig"] * s_ref[k][0]), k
    err = engine.rel_error(cores, dims, used, A)
    err_ref = ttsvd.rel_error(An, ttsvd.tt_full(c_ref, dims, used))
    assert abs(err - err_ref) <= tol["err"] + 2e-2 * err_ref, (err, err_ref)
    for k in range(len(dims) - 1):
        g, shape = cores[k]
        assert rel(fmat(g, *shape), c_ref[k]) < 10 * tol["vec"], k


@pytest.mark.parametrize("dims,ranks,s,q,mode", [
    ([32, 32, 24, 20], [1, 32, 40, 12, 1], 10, 0, "fixed"),
    ([32, 32, 24, 20], [1, 32, 40, 12, 1], 10, 1, "fixed"),
    ([32, 16, 32, 8, 6], [1, 32, 48, 30, 6, 1], 8, 0, "fixed"),
])
def test_tt_svd_fused_first_pass(engine, dims, ranks, s, q, mode):
    """BF16 sweeps whose first read of the tensor serves both the exact 32-row first step (its Gram)
    and the second step's 64-column sketch: the same results as the oracle, which sketches the
    projected remainder, and one pass fewer (step 0 counts the shared pass, step 1 only its
    projections)."""
    A = engine.generate(dims, ranks, std=1.0, noise_rel=1e-6, seed=2024)
    cores, sig, rep = engine.tt_svd(A, dims, ranks, s=s, q=q, precision="bf16", rank_mode=mode)
    An = host(A).astype(np.float64)
    c_ref, s_ref, p_ref, used = ttsvd.tt_svd(An, dims, ranks, s=s, q=q)
    tol = TOL["bf16"]
    assert rep["ranks"] == used
    assert rep["passes"][0] == 1 and rep["passes"][1] == 1 + 2 * q, rep["passes"]
    for k in range(len(dims) - 1):
        r = ranks[k + 1]
        assert np.allclose(sig[k][:r], s_ref[k][:r], rtol=tol["sig"], atol=tol["sig"] * s_ref[k][0]), k
    err = engine.rel_error(cores, dims, used, A)
    err_ref = ttsvd.rel_error(An, ttsvd.tt_full(c_ref, dims, used))
    assert abs(err - err_ref) <= tol["err"] + 2e-2 * err_ref, (err, err_ref)
    for k in range(len(dims) - 1):
        g, shape = cores[k]
        assert rel(fmat(g, *shape), c_ref[k]) < 10 * tol["vec"], k


@pytest.mark.parametrize("dims,ranks,s,q", [
    ([8, 9, 10, 11, 6], [1, 4, 6, 6, 5, 1], 4, 1),
    ([32, 32, 24, 20], [1, 32, 40, 12, 1], 10, 0),       # fused first pass (three-product Gram)
    ([32, 16, 32, 8, 6], [1, 32, 48, 30, 6, 1], 8, 0),    # fused first pass, five modes
])
def test_tt_svd_tail_x3_matches_oracle(engine, dims, ranks, s, q):
    """TTX_FLAG_TAIL_X3 (three bf16 products in the randomized steps' projection): the oracle's ranks,
    singular values, cores and error to the BF16 tolerances, and the error within 2e-6 of the
    four-product sweep's."""
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    A = engine.generate(dims, ranks, std=1.0, noise_rel=1e-6, seed=1134)
    cores, sig, rep = engine.tt_svd(A, dims, ranks, s=s, q=q, precision="bf16", flags=L.FLAG_TAIL_X3)
    c4, _, rep4 = engine.tt_svd(A, dims, ranks, s=s, q=q, precision="bf16")
    An = host(A).astype(np.float64)
    c_ref, s_ref, p_ref, used = ttsvd.tt_svd(An, dims, ranks, s=s, q=q)
    tol = TOL["bf16"]
    assert rep["ranks"] == used == rep4["ranks"]
    for k in range(len(dims) - 1):
        r = ranks[k + 1]
        assert np.allclose(sig[k][:r], s_ref[k][:r], rtol=tol["sig"], atol=tol["sig"] * s_ref[k][0]), k
    err = engine.rel_error(cores, dims, used, A)
    err4 = engine.rel_error(c4, dims, used, A)
    err_ref = ttsvd.rel_error(An, ttsvd.tt_full(c_ref, dims, used))
    assert abs(err - err_ref) <= tol["err"] + 2e-2 * err_ref, (err, err_ref)
    assert abs(err - err4) <= 2e-6, (err, err4)
    for k in range(len(dims) - 1):
        g, shape = cores[k]
        assert rel(fmat(g, *shape), c_ref[k]) < 10 * tol["vec"], k


# ------------------------------------------------------------------------------------------------
# adaptive range finder: the engine takes the oracle's decisions (widths, power iterations, checks)
# ------------------------------------------------------------------------------------------------
def _flags():
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    return L.FLAG_ADAPTIVE


def _check_adaptive(rep, infos, p_ref):
    for k, inf in enumerate(infos):
        if p_ref[k] != "rsvd":
            assert rep["range_checks"][k] == 0, k
            continue
        assert rep["path"][k] == 2, k
        assert rep["ell"][k] == inf["ell"], (k, rep["ell"], inf)
        assert rep["power_iters"][k] == inf["power_iters"], (k, rep["power_iters"], inf)
        assert rep["range_checks"][k] == inf["checks"], (k, rep["range_checks"], inf)
        got, want = rep["range_resid"][k] ** 2, inf["rel2"]
        assert abs(got - want) <= 0.05 * want + 1e-10, (k, got, want)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_adaptive_fixed_rank_stops_at_q0(engine, prec):
    dims, ranks = [6, 8, 8, 8, 8], [1, 5, 6, 6, 5, 1]
    A = engine.generate(dims, ranks, std=1.0, noise_rel=1e-6, seed=21)
    cores, sig, rep = engine.tt_svd(A, dims, ranks, s=6, q=2, precision=prec, flags=_flags())
    An = host(A).astype(np.float64)
    infos = []
    c_ref, s_ref, p_ref, used = ttsvd.tt_svd(An, dims, ranks, s=6, q=2, adaptive=True, infos=infos)
    assert rep["ranks"] == used
    _check_adaptive(rep, infos, p_ref)
    for k, p in enumerate(p_ref):
        if p == "rsvd":
            assert rep["power_iters"][k] == 0 and rep["passes"][k] == 2, k
    tol = TOL[prec]
    for k in range(len(dims) - 1):
        r = ranks[k + 1]
        assert np.allclose(sig[k][:r], s_ref[k][:r], rtol=tol["sig"], atol=tol["sig"] * s_ref[k][0]), k
    err = engine.rel_error(cores, dims, used, A)
    err_ref = ttsvd.rel_error(An, ttsvd.tt_full(c_ref, dims, used))
    assert abs(err - err_ref) <= tol["err"] + 2e-2 * err_ref, (err, err_ref)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_adaptive_delta_grows_basis(engine, prec):
    # true ranks 20 under a ceiling of 32: the 16-column basis fails its check, 32 columns pass
    dims, true = [4, 12, 12, 12, 4], [1, 4, 20, 20, 4, 1]
    A = engine.generate(dims, true, std=1.0, noise_rel=1e-6, seed=22)
    ceil = [1, 4, 32, 32, 4, 1]
    cores, sig, rep = engine.tt_svd(A, dims, ceil, s=10, q=2, precision=prec, rank_mode="delta", eps=1e-3,
                                    flags=_flags())
    An = host(A).astype(np.float64)
    infos = []
    c_ref, s_ref, p_ref, used = ttsvd.tt_svd(An, dims, ceil, s=10, q=2, rank_mode="delta", eps=1e-3,
                                             adaptive=True, infos=infos)
    assert used == true
    assert rep["ranks"] == used
    _check_adaptive(rep, infos, p_ref)
    assert [inf["ell"] for inf in infos if inf] == [32, 32]
    err = engine.rel_error(cores, dims, used, A)
    assert err < 1e-3


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("case", ["one_iteration", "ceiling"])
def test_adaptive_power_iterations(engine, prec, case):
    # 8 strong directions over a flat tail (one step, d = 2): q = 0 leaves a tail-sized residual
    rng = np.random.default_rng(9)
    m, n = 300, 800
    sig = np.concatenate([np.ones(8), np.full(200, 3e-3)])
    U = np.linalg.qr(rng.standard_normal((m, len(sig))))[0]
    V = np.linalg.qr(rng.standard_normal((n, len(sig))))[0]
    A64 = (U * sig) @ V.T
    A32 = A64.astype(np.float32)
    An = A32.astype(np.float64)
    r0 = ttsvd.adaptive_range(An, [10], 8, 0, 0.0, 1234, 0)[3]
    r1 = ttsvd.adaptive_range(An, [10], 8, 1, 0.0, 1234, 0)[3]
    assert r1 < 0.5 * r0
    tol2 = np.sqrt(r0 * r1) if case == "one_iteration" else 1e-3 * r1
    eps = float(np.sqrt(4.0 * tol2))     # d = 2: tol2 = eps^2 / 4
    cores, s_gpu, rep = engine.tt_svd(dev(A32), [m, n], [1, 8, 1], s=2, q=3, precision=prec, eps=eps,
                                      flags=_flags())
    infos = []
    c_ref, s_ref, p_ref, used = ttsvd.tt_svd(An, [m, n], [1, 8, 1], s=2, q=3, eps=eps, adaptive=True,
                                             infos=infos)
    want_it = 1 if case == "one_iteration" else 3
    assert infos[0]["power_iters"] == want_it
    _check_adaptive(rep, infos, p_ref)
    assert rep["passes"][0] == 2 + 2 * want_it
    tol = TOL[prec]
    assert np.allclose(s_gpu[0][:8], s_ref[0][:8], rtol=tol["sig"], atol=tol["sig"] * s_ref[0][0])


# ------------------------------------------------------------------------------------------------
# FP16 range finder with power-of-two operand scaling
# ------------------------------------------------------------------------------------------------
def test_fp16_tt_svd_close_to_oracle(engine):
    # binary16 sketch operands (11 significant bits) perturb the range basis by about 2^-11 relative;
    # the projection and remainder stay 4-term bf16 accurate, so the reconstruction error stays within
    # 2e-3 of the FP64 oracle's
    dims, true = [8, 9, 10, 11, 6], [1, 4, 6, 6, 5, 1]
    A = engine.generate(dims, true, std=1.0, noise_rel=1e-6, seed=1134)
    cores, sig, rep = engine.tt_svd(A, dims, true, s=4, q=1, precision="fp16")
    An = host(A).astype(np.float64)
    c_ref, s_ref, p_ref, used = ttsvd.tt_svd(An, dims, true, s=4, q=1)
    assert rep["ranks"] == used
    err = engine.rel_error(cores, dims, used, A)
    err_ref = ttsvd.rel_error(An, ttsvd.tt_full(c_ref, dims, used))
    assert err <= err_ref + 2e-3, (err, err_ref)
    for k in range(len(dims) - 1):
        r = true[k + 1]
        assert np.allclose(sig[k][:r], s_ref[k][:r], rtol=1e-3, atol=1e-3 * s_ref[k][0]), k


def test_fp16_power_of_two_scaling_is_exact(engine):
    # entries of 1e6 would overflow binary16 (max 65504); with the power-of-two operand scales the
    # decomposition of 2^20 A is that of A scaled by 2^20 (every other step is scale-invariant)
    import torch
    dims, true = [6, 20, 20, 20], [1, 6, 8, 6, 1]
    A = engine.generate(dims, true, std=1.0, noise_rel=1e-4, seed=77)
    big = (A * float(2 ** 20)).contiguous()
    assert float(big.abs().max()) > 65504.0
    c1, s1, r1 = engine.tt_svd(A, dims, true, s=10, q=1, precision="fp16")
    c2, s2, r2 = engine.tt_svd(big, dims, true, s=10, q=1, precision="fp16")
    assert r1["path"][1] == 2    # the middle step is randomized (fp16 range finder)
    assert np.all(np.isfinite(s2))
    assert np.allclose(s2, s1 * 2.0 ** 20, rtol=1e-6, atol=0)
    for k in range(len(dims)):
        g1, g2 = host(c1[k][0]), host(c2[k][0])
        if k == len(dims) - 1:
            g2 = g2 / 2.0 ** 20
        assert np.max(np.abs(g1 - g2)) <= 1e-6 * np.max(np.abs(g1)), k
    err = engine.rel_error(c2, dims, true, big)
    assert err < 1e-3


def test_permute_231_bit_exact_at_2_pow_31(engine):
    """The TR first-step relayout at 32^6 with r0 r1 = 4 x 16: 64 x 32^5 = 2^31 elements (8.6 GB in, 8.6 GB
    out), past the int32 index range where the reference's tensorToMode231 overflows
    (trdec6_singleGPU.cu:51).  Every element carries its own source index as its bit pattern."""
    import torch
    r0, r1, n = 4, 16, 32 ** 5
    total = r0 * r1 * n
    assert total == 2 ** 31
    x = torch.empty(total, dtype=torch.int32, device="cuda")
    step = 1 << 28
    for c in range(0, total, step):
        x[c:c + step] = torch.arange(c, c + step, dtype=torch.int64, device="cuda").to(torch.int32)
    out = engine.permute_231(x.view(torch.float32), r0, r1, n)
    del x
    o = out.view(torch.int32)
    # out position beta + r1 (j + n alpha) holds source index alpha + r0 (beta + r1 j)
    beta = torch.arange(r1, dtype=torch.int64, device="cuda")
    rows = 1 << 22
    for alpha in range(r0):
        sl = o[alpha * r1 * n:(alpha + 1) * r1 * n].view(n, r1)
        for j0 in range(0, n, rows):
            j = torch.arange(j0, min(n, j0 + rows), dtype=torch.int64, device="cuda")
            want = alpha + r0 * beta[None, :] + (r0 * r1) * j[:, None]
            assert torch.equal(sl[j0:j0 + rows].to(torch.int64), want), (alpha, j0)
    torch.cuda.synchronize()


def test_device_args_refuse_nd_tensors(engine):
    """Device tensors of the decomposition entry points hold column-major data and must be flat; a
    C-contiguous N-d tensor would be read transposed, so it is refused before any call."""
    import torch
    from tensor_train_ring_decompositions_hpc_gpu_amd.api import _dev_ptr
    A = torch.zeros(4, 5, 6, device="cuda", dtype=torch.float32)
    with pytest.raises(TypeError, match="flat"):
        engine.tt_svd(A, [4, 5, 6], [1, 2, 2, 1])
    assert _dev_ptr(A.reshape(-1)) == A.data_ptr()
    assert _dev_ptr(A.reshape(20, 6), flat=False) == A.data_ptr()   # row-major batches of the rank model


def test_jacobi_sweep_cap_reports_no_convergence(monkeypatch):
    """The eigensolver's sweep cap is a failure, not a silent result: with the cap forced down to one
    sweep from the host (TTX_JACOBI_MAX_SWEEPS, read at handle creation) a decomposition whose Gram
    needs several sweeps returns TTX_ERR_NO_CONVERGE, through the sweep and through the single-step
    entry point; a fresh handle with the default cap converges on the same input."""
    import torch
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    from tensor_train_ring_decompositions_hpc_gpu_amd.api import Engine
    monkeypatch.setenv("TTX_JACOBI_MAX_SWEEPS", "1")
    eng = Engine()
    try:
        dims, ranks = [12, 10, 9, 8], [1, 6, 6, 5, 1]
        A = eng.generate(dims, ranks, std=1.0, noise_rel=1e-6, seed=77)
        with pytest.raises(L.TtxError) as ei:
            eng.tt_svd(A, dims, ranks, s=4, q=1, precision="fp32")
        assert ei.value.status == L.ERR_NO_CONVERGE and "sweeps" in str(ei.value)
        with pytest.raises(L.TtxError) as ei:
            eng.step(A, 12, 720, 6, s=4, q=1, precision="fp32")
        assert ei.value.status == L.ERR_NO_CONVERGE
    finally:
        torch.cuda.synchronize()
        eng.close()
    monkeypatch.delenv("TTX_JACOBI_MAX_SWEEPS")
    eng = Engine()
    try:
        _, _, rep = eng.tt_svd(A, dims, ranks, s=4, q=1, precision="fp32")
        assert max(rep["jacobi_sweeps"]) > 1
    finally:
        torch.cuda.synchronize()
        eng.close()


def _tr_case(engine, dims, true, ranks, s, q, prec, std, seed=1134, flags=0):
    A = engine.generate(dims, true, std=std, noise_rel=1e-6, seed=seed, tr=True)
    cores, sig, rep = engine.tr_svd(A, dims, ranks, s=s, q=q, precision=prec, flags=flags)
    An = host(A).astype(np.float64)
    c_ref, s_ref, p_ref, used = ttsvd.tr_svd(An, dims, ranks, s=s, q=q)
    err = engine.rel_error(cores, dims, ranks, A, tr=True)
    err_ref = ttsvd.rel_error(An, ttsvd.tr_full(c_ref, dims, ranks))
    return cores, sig, rep, c_ref, s_ref, p_ref, err, err_ref


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_tr_svd_zero_padded_first_core(engine, prec):
    """r0 r1 = 3 x 4 = 12 columns on a 6-row first unfolding: the first core keeps its 12 columns, the
    last 6 zero (rank(A_1) <= 6), and the 231 relayout still sees r0 x r1 rows."""
    dims, true, ranks = [6, 8, 8, 7], [3, 4, 4, 3, 3], [3, 4, 12, 9, 3]
    cores, sig, rep, c_ref, s_ref, p_ref, err, err_ref = _tr_case(engine, dims, true, ranks, 4, 1, prec, 0.7)
    tol = TOL[prec]
    g0 = fmat(cores[0][0], 6, 12)
    assert np.all(g0[:, 6:] == 0.0)
    assert np.allclose(sig[0][:6], s_ref[0][:6], rtol=tol["sig"], atol=tol["sig"] * s_ref[0][0])
    for k in range(1, len(dims) - 1):
        r = ranks[k + 1]
        assert np.allclose(sig[k][:r], s_ref[k][:r], rtol=tol["sig"], atol=tol["sig"] * s_ref[k][0]), k
    assert abs(err - err_ref) <= tol["err"] + 2e-2 * err_ref, (err, err_ref)


def test_tr_svd_fp16_well_posed_ring(engine):
    """A ring of ranks 2 with inner ranks r0^2 r_k = 8 is reproduced exactly by the TR sweep (the oracle
    reaches the 1e-6 noise floor); with the FP16 range finder the error stays within 5e-4 of it."""
    dims, true, ranks = [8] * 6, [2] * 7, [2, 2, 8, 8, 8, 8, 2]
    cores, sig, rep, c_ref, s_ref, p_ref, err, err_ref = _tr_case(engine, dims, true, ranks, 10, 2, "fp16", 0.5)
    assert err_ref < 1e-5
    assert err <= err_ref + 5e-4, (err, err_ref)
    for k in range(len(dims) - 1):
        r = min(ranks[0] * ranks[1] if k == 0 else ranks[k + 1], len(s_ref[k]))
        assert np.allclose(sig[k][:r], s_ref[k][:r], rtol=1e-3, atol=1e-3 * s_ref[k][0]), k


@pytest.mark.parametrize("prec", ["fp16", "fp32"])
def test_tr_svd_reduced_cfg3_schedule(engine, prec):
    """The TR workload's schedule at 12^6: ring cores 16 x 12 x 16 N(0, 1/4), first-step split
    r0 r1 = 4 x 16 (64 columns on a 12-row unfolding, zero-padded), inner ranks 16, s = 10, q = 2,
    with the adaptive check as the bench runs it.  The split keeps a ring rank of 4 where the data has
    16, so most of the energy is truncated by construction: the FP64 oracle's error for this schedule
    is 0.9776 at 12^6 (0.842 at 6^6, 0.932 at 8^6, 0.987 at 16^6, rising towards 1 with the mode size;
    the engine reports 0.997 at 32^6).  The engine must reproduce the oracle's error and spectra."""
    from tensor_train_ring_decompositions_hpc_gpu_amd import _lib as L
    dims, true, ranks = [12] * 6, [16] * 7, [4, 16, 16, 16, 16, 16, 4]
    cores, sig, rep, c_ref, s_ref, p_ref, err, err_ref = _tr_case(engine, dims, true, ranks, 10, 2, prec, 0.5,
                                                                  flags=L.FLAG_ADAPTIVE)
    assert abs(err_ref - 0.9776) < 5e-3, err_ref
    assert rep["path"] == [0, 2, 2, 2, 2] and rep["power_iters"][1:] == [2, 2, 2, 2], rep
    assert abs(err - err_ref) <= 1e-3, (err, err_ref)
    assert np.allclose(sig[0][:12], s_ref[0][:12], rtol=1e-4, atol=1e-4 * s_ref[0][0])
    for k in range(1, len(dims) - 1):
        assert np.allclose(sig[k][:16], s_ref[k][:16], rtol=1e-3, atol=1e-3 * s_ref[k][0]), k
