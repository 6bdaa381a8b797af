// Host-side launchers for every ttx kernel.  All launches are asynchronous on the given stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ttx {

int num_cus();

// ---- streaming passes (ttx_stream.hip) ---------------------------------------------------------
struct SketchPlan {
  int RB = 1, CG = 1, ysplit = 1, grid_x = 1, kp = 32, lds_bytes = 0;
  int64_t chunks_per_wg = 1, slabs = 1;
  bool plain = false;   // the register-buffered kernel only (the one batches group)
};
SketchPlan plan_sketch(int m, int64_t n, int kp, int wgs_hint);
// a batch's plan: the same tiling, cpw chunks per WG (partial slabs of about kp / (16 cpw) of A's size),
// the register-buffered kernel
SketchPlan plan_sketch_batch(int m, int64_t n, int kp, int64_t cpw);
// Y partials: part[slab][m][kp].  X == nullptr selects the in-kernel Philox sketch (seed, stream, ell).
// FP16 arithmetic: scales (optional) holds power-of-two operand scales, [0] for A and [1 + c] for
// column c of X, undone on the partial sums.
void launch_sketch(int ar, const SketchPlan& p, const float* A, int64_t lda, int m, int64_t n, const float* X,
                   int64_t ldx, uint64_t seed, uint32_t stream, int ell, float* part, hipStream_t st,
                   const float* scales = nullptr);
// The fused first pass of a TT sweep: the BF16X3 Omega sketch of A (m x n, kp = 64 columns) as
// launch_sketch, plus the Gram of the 32-row view of the same memory (32 x (m / 32) n) into
// sketch_gram0_partials(p) FP64 32 x 32 partials.  sketch_gram0_ok: the shapes it takes.
bool sketch_gram0_ok(const SketchPlan& p, int m, int64_t lda, int kp);
int64_t sketch_gram0_partials(const SketchPlan& p);
// omega: Omega precomputed by launch_omega_chunks for (seed, stream, n, ell) (the pass then spends no
// arithmetic on it), or nullptr (generated in the pass).  gram3 (with omega): the Gram with three bf16
// products (no mid mid term) instead of four.
void launch_sketch_gram0(const SketchPlan& p, const float* A, int64_t lda, int m, int64_t n, uint64_t seed,
                         uint32_t stream, int ell, float* part, double* gpart, hipStream_t st,
                         const void* omega = nullptr, bool gram3 = false);
// Omega (kp = 64 columns, the first ell nonzero) of an n-column sketch in 2 KB chunks of 16 rows (bf16,
// exact): omega_chunks_bytes(n) bytes at out
int64_t omega_chunks_bytes(int64_t n);
void launch_omega_chunks(uint64_t seed, uint32_t stream, int64_t n, int ell, void* out, hipStream_t st);
// scales[0] from sum_i<r sig_prev[i]^2 / numel (RMS entry to [8, 16)); sig_prev == nullptr gives 1
void launch_fp16_ascale(const float* sig_prev, int r_prev, double numel, float* scales, hipStream_t st);
// scales[1 + c] from max_j |X[j][c]| (to [128, 256)); mx: kp zeroed words, zeroed again afterwards
void launch_fp16_xscale(const float* X, int64_t n, int64_t ldx, int kp, unsigned* mx, float* scales, hipStream_t st);
// scales[1 + c] from the column maxima mx[c] (float bits) alone; mx zeroed again afterwards
void launch_fp16_colscale(unsigned* mx, int kp, float* scales, hipStream_t st);
// D = Q^T A; mode 0 writes Z (n x ldo row-major: kp columns, or the first r_out when r_out > 0, a multiple
// of 4), mode 1 writes r_out x n column-major, mode 2 writes the TR first-step remainder straight into the
// 231 layout: row c = a + r0 b (b < r1c) of column j to out[b + r1c (j + n a)], with r0 = ldo and
// r1c = r_out (rows c >= r0 r1c are not written).
// gpart != nullptr: also the Gram D D^T, first kg rows / columns, one FP64 partial per WG
// (proj_grid(n) partials).
int64_t proj_grid(int64_t n);
// FP16 arithmetic: scales[0] (optional) is a power-of-two scale of A, undone on the products.
void launch_proj(int ar, int kp, const float* A, int64_t lda, int m, int64_t n, const float* Qc, int64_t ldq,
                 int mode, float* out, int64_t ldo, int r_out, hipStream_t st, double* gpart = nullptr, int kg = 0,
                 const float* scales = nullptr);
// D = Q^T A for a short unfolding (m <= 64 rows; the exact step's G^T A): mode 1 (r_out x n column-major)
// or mode 2 (the TR first step's 231 layout, as launch_proj); arithmetic AR_FP32 or AR_BF16X4 (any other
// takes AR_BF16X4); kp = 32 or 64 columns of Qc (ldq); grid 0: proj_short_grid(n, false)
int64_t proj_short_grid(int64_t n, bool batch);
void launch_proj_short(int ar, int kp, const float* A, int64_t lda, int m, int64_t n, const float* Qc, int64_t ldq,
                       int mode, float* out, int64_t ldo, int r_out, int64_t grid, hipStream_t st);
int64_t gram_small_grid(int64_t n);
// partial A A^T: part[grid][m][m] (FP64); absmax (optional) receives max |A| as float bits
void launch_gram_small(int ar, const float* A, int64_t lda, int m, int64_t n, double* part, int64_t grid,
                       float* absmax, hipStream_t st);
void launch_reduce_f32(const float* part, int64_t S, int64_t count, float* out, hipStream_t st);
void launch_reduce_f64(const double* part, int64_t S, int64_t count, double* out, hipStream_t st);

// ---- FP64 small dense kernels (ttx_dense.hip) --------------------------------------------------
// Gram of a tall matrix X (rows x k, row stride ldx): part[grid][k][k] FP64.
int64_t gram_tall_grid(int64_t rows);
void launch_gram_tall_f32(const float* X, int64_t ldx, int64_t rows, int k, double* part, int64_t grid,
                          hipStream_t st);
void launch_gram_tall_f64(const double* X, int64_t ldx, int64_t rows, int k, double* part, int64_t grid,
                          hipStream_t st);
// number of 64-row blocks of the CholeskyQR panel kernels
int64_t panel_blocks(int64_t rows);
// Sum S Gram partials, Cholesky of G (+ shift) and the explicit inverse of the upper factor:
// G + s I = R^T R, Rinv = R^-1.  shift_mode 1 adds s = 11 (m k + k (k+1)) u trace(G); pivots that are
// not positive zero their row and column of Rinv and count into status[0].  pivots (optional, k doubles)
// receives the shifted pivots R_ii^2.
// Rinv null: the factorisation only (pivots and status)
// kinv (0: k): only the leading kinv x kinv block of Rinv is computed (the rest of Rinv is unspecified).
void launch_chol_inv(const double* part, int64_t S, int k, int64_t m, int shift_mode, double* Rinv, int* status,
                     hipStream_t st, double* pivots = nullptr, int kinv = 0);
// W_out = W_in Rinv (rows x k) per 64-row block with the Gram partial of W_out (part[block][k][k]).
// ldr: row stride of Rinv (0: k), so the leading block of a wider factor's inverse can be used.
void launch_apply_gram_f32(const float* Win, int64_t ldi, int64_t rows, int k, const double* Rinv, double* Wout,
                           double* part, hipStream_t st, int ldr = 0);
void launch_apply_gram_f64(const double* Win, int64_t rows, int k, const double* Rinv, double* Wout, double* part,
                           hipStream_t st);
// final round: Qc (column-major kp x m, FP32) = W_in Rinv, columns k..kp-1 zeroed
void launch_apply_out(const double* Win, int64_t rows, int k, int kp, const double* Rinv, float* Qc, int64_t ldq,
                      hipStream_t st);
// final round as one Newton-Schulz step: Qc = W_in (3 I - G) / 2 with G the sum of S Gram partials of W_in
void launch_apply_ns_out(const double* Win, int64_t rows, int k, int kp, const double* gpart, int64_t S, float* Qc,
                         int64_t ldq, hipStream_t st);
// symmetric eigen-decomposition (cyclic Jacobi, FP64): lam descending, V row-major k x k (columns are
// eigenvectors); status[1] = sweeps, status[2] = 1 when max_sweeps ran out before convergence;
// keys[0..63] are zeroed for the following assembly
constexpr int kJacobiMaxSweeps = 30;
void launch_syevj(const double* G, int k, double* lam, double* V, int* status, unsigned long long* keys,
                  hipStream_t st, int max_sweeps = kJacobiMaxSweeps);
// core assembly after the eigen-decomposition of the projected Gram:
//   Gc = Q V[:, :r] (m x r column-major, FP32) with the sign of each column fixed so that its largest
//   |entry| is positive (first index on ties); Ur = V[:, :r] diag(sign) (row-major FP32, round4(k) rows
//   of next_from_z_ld(r) floats, zero beyond k rows and r columns);
//   Qn = Gc padded with zero columns to kpn (column-major kpn x m) for a following projection;
//   sigma = sqrt(max(lam, 0)) for the first k values.  Qc == nullptr means Q = I (exact path).
//   keys: 64 zeroed 64-bit words (launch_syevj zeroes them).
void launch_assemble(const float* Qc, int64_t ldq, int m, int k, const double* lam, const double* V, int r,
                     float* Gc, float* Ur, float* Qn, int kpn, float* sigma, unsigned long long* keys,
                     hipStream_t st);
// A_next[c + r j] = sum_a Ur[a][c] Z[j][a]  (c < r, j < n); Z row-major with stride ldz (a multiple of 4,
// 16-byte aligned rows, columns k .. round_up(k, 4) - 1 readable); Ur as launch_assemble writes it.
// batch: one output row per lane for every n
int next_from_z_ld(int r);
void launch_next_from_z(const float* Z, int64_t ldz, int64_t n, int k, const float* Ur, int r, float* out,
                        hipStream_t st, bool batch = false);
// pending full-rank left factor P = I_N (x) G^T of a fused step (G: ms x ms column-major core):
// mode 0: Y = P X (X, Y row-major m x kp); mode 1: Y = P^T X (X, Y column-major kp x m); mode 2: Y = P X
// (X, Y column-major kp x m)
void launch_blockmix(const float* G, int ms, int m, int kp, int mode, const float* X, float* Y, hipStream_t st);
// Ca = P Cs P^T (m x m, FP64)
void launch_gram_mix(const float* G, int ms, int m, const double* Cs, double* Ca, hipStream_t st);
// a-posteriori range check: out[2 b] = ||(I - Q Q^T) Yt||^2 and out[2 b + 1] = ||Yt||^2 over 64-row block
// b (panel_blocks(m) blocks); Q = first k columns of Qc (column-major, ldq), Yt m x p row-major (p <= 16);
// part: panel_blocks(m) * k * p doubles of scratch
void launch_test_resid(const float* Qc, int64_t ldq, int m, int k, const float* Yt, int p, double* part, double* out,
                       hipStream_t st);
// Yt (m x p row-major) <- columns c0 .. c0+p-1 of Y (m x kpy row-major)
void launch_copy_block(const float* Y, int kpy, int m, int c0, int p, float* Yt, hipStream_t st);
// TR core with zero rows restored: full[(b + r1 i) + r1 N c] = b < r1c ? compact[(b + r1c i) + r1c N c] : 0
// for b < r1, i < N, c < r (both column-major (r1 N) x r and (r1c N) x r)
void launch_expand_rows(const float* compact, int r1c, int r1, int64_t N, int r, float* full, hipStream_t st);
// Y (m x kp row-major) <- A (m x n column-major, n <= kp), zero columns n..kp-1
void launch_copy_cols(const float* A, int64_t lda, int m, int n, float* Y, int kp, hipStream_t st);

// ---- generators / reconstruction (ttx_rng.hip, ttx_gemm.hip) -----------------------------------
void launch_fill_normal(float* out, int64_t count, uint64_t seed, uint32_t stream, float scale, hipStream_t st);
// zero bytes (a multiple of 4) with a kernel
void launch_zero(void* p, int64_t bytes, hipStream_t st);
void launch_add_noise(float* A, int64_t count, uint64_t seed, uint32_t stream, float coef, hipStream_t st);
void launch_sumsq(const float* A, int64_t count, double* out2 /* [1] */, double* part, int64_t grid, hipStream_t st);
int64_t sumsq_grid(int64_t count);
void launch_omega(float* out, int64_t n, int ell, uint64_t seed, uint32_t stream, hipStream_t st);
// C = alpha A B + beta C (column-major, FP32); resid != nullptr: no C write, accumulate
// sum (X - A B)^2 and sum X^2 per block into resid[2 * block] (X with leading dimension ldc)
void launch_gemm(int64_t M, int64_t N, int64_t K, float alpha, const float* A, int64_t lda, const float* B,
                 int64_t ldb, float beta, float* C, int64_t ldc, hipStream_t st);
int64_t gemm_resid_blocks(int64_t M, int64_t N);
void launch_gemm_resid(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda, const float* B, int64_t ldb,
                       const float* X, int64_t ldx, double* resid, hipStream_t st);
// generic permutation of a column-major tensor with up to 6 modes: out[perm] = in
void launch_permute(const float* in, float* out, int nd, const int64_t* dims, const int* perm, hipStream_t st);
// (r0, r1, n) -> (r1, n, r0), column-major (trdec6_singleGPU.cu:49-58)
void launch_permute_231(const float* in, float* out, int64_t r0, int64_t r1, int64_t n, hipStream_t st);

// ---- rank model (ttx_mlp.hip) -----------------------------------------------------------------
void mlp_init(float* params, int in, int hid, int out, uint64_t seed, hipStream_t st);
// features from the probe's shifted Cholesky pivots d[0..kp-1] (= R_ii^2)
void mlp_features(const double* d, int kp, double rel_tol, double m, double n, float* x, int in, hipStream_t st);
// features + forward of one sample in one workgroup (same arithmetic as mlp_features + mlp_forward)
void mlp_predict(const float* P, const double* d, int kp, double rel_tol, double m, double n, float* y,
                 hipStream_t st);
void mlp_forward(const float* P, int in, int hid, int out, const float* X, int B, float* Z1, float* H1, float* Z2,
                 float* H2, float* Y, hipStream_t st);
void mlp_backward(const float* P, int in, int hid, int out, const float* X, const float* T, int B, const float* Z1,
                  const float* H1, const float* Z2, const float* H2, const float* Y, float* dY, float* dH,
                  float* dZ, float* grad, double* part, hipStream_t st);
int mlp_loss_blocks(int64_t count);
void mlp_adam(float* P, const float* g, float* m1, float* m2, int count, float lr, int t, hipStream_t st);

// ---- grouped launches (batches) ---------------------------------------------------------------
// One launch runs one kernel over many jobs: hardware block b runs block map[b].y of job map[b].x.
// Each job descriptor starts with its grid (gx x gy, the grid its plain launch would have) and
// carries that launch's arguments, so a job's results are those of its plain launch, bit for bit.
struct GroupLaunch {
  const void* jobs = nullptr;   // device array of descriptors
  const int2* map = nullptr;    // device array, one entry per block
  int blocks = 0, threads = 0, lds = 0;
};
struct JobGramSmall { int gx, gy; const float* A; int64_t lda; int m; int64_t n; int64_t cpw; double* part; };
struct JobReduce { int gx, gy; const void* part; int64_t S; int64_t count; void* out; };
struct JobSketch {
  int gx, gy; const float* A; int64_t lda; int m; int64_t n; const float* X; int64_t ldx; uint64_t seed;
  uint32_t stream; int ell; int RB, CG; int64_t cpw; float* part;
};
struct JobProj {
  int gx, gy; const float* A; int64_t lda; int m; int64_t n; const float* Qc; int64_t ldq; int mode; float* out;
  int64_t ldo; int r_out; int64_t blocks; double* gpart; int kg;
};
struct JobGramRows { int gx, gy; const void* X; int64_t ldx; int64_t rows; int k; int64_t rpw; double* part; };
struct JobChol { int gx, gy; const double* part; int64_t S; int k; int64_t m; int shift; double* Rinv; int* status; };
struct JobApply {
  int gx, gy; const void* Win; int64_t ldi; int64_t rows; int k; int kp; const double* Rinv; double* Wd; float* Qc;
  int64_t ldq; double* part; const double* gpart; int64_t S; int ldr;
};
struct JobSyevj { int gx, gy; const double* G; int k; double* lam; double* V; int* status; unsigned long long* keys; int nt; };
struct JobAssemble {
  int gx, gy; const float* Qc; int64_t ldq; int m; int k; const double* V; const double* lam; int r; float* Gc;
  float* Ur; float* Qn; int kpn; float* sigma; unsigned long long* keys;
};
struct JobNext { int gx, gy; const float* Z; int64_t ldz; int64_t n; int k; const float* Ur; int r; float* out; int rm; int cols; };
struct JobCopyCols { int gx, gy; const float* A; int64_t lda; int m; int n; float* Y; int kp; };

// grids the plain launches use, so a batch plan can build the same jobs
int64_t proj_blocks(int64_t n);
int syevj_threads(int k);
// batch: one row per lane whatever n (the whole batch fills the machine), at most 4096 workgroups
int64_t next_from_z_grid(int64_t n, int r, bool* cols, bool batch = false);
int64_t copy_cols_grid(int m, int kp);
// plain launches with an explicit grid (the batch plans' grids; grid 0: the default)
// colmax (FP16 arithmetic, mode 0 only): also atomicMax of max_j |Z[j][c]| (float bits) into colmax[c],
// the input of launch_fp16_colscale (the column scales of Z without launch_fp16_xscale's pass over it)
void launch_proj_grid(int ar, int kp, const float* A, int64_t lda, int m, int64_t n, const float* Qc, int64_t ldq,
                      int mode, float* out, int64_t ldo, int r_out, int64_t grid, hipStream_t st, double* gpart, int kg,
                      const float* scales = nullptr, unsigned* colmax = nullptr);
// grouped launchers (ar: the Arith of every job; variant: the template choice the jobs share)
void grp_gram_small(int ar, int mt, const GroupLaunch& g, hipStream_t st);
void grp_reduce(bool f64, const GroupLaunch& g, hipStream_t st);
void grp_sketch(int ar, int lt, bool omega, int fq, const GroupLaunch& g, hipStream_t st);
void grp_proj(int ar, int lt, bool gram, const GroupLaunch& g, hipStream_t st);
void grp_proj_short(int ar, int lt, int nk, const GroupLaunch& g, hipStream_t st);
// kc: the widest job's k (the LDS is sized for 32, 48 or 64 by grp_kcap)
int grp_kcap(int k);
void grp_gram_rows(bool f64, int kc, const GroupLaunch& g, hipStream_t st);
void grp_chol(const GroupLaunch& g, hipStream_t st);
void grp_apply(int kind, int kc, const GroupLaunch& g, hipStream_t st);   // kind 0: float in, 1: double in, 2: Newton-Schulz out
void grp_syevj(const GroupLaunch& g, hipStream_t st, int max_sweeps);
void grp_assemble(const GroupLaunch& g, hipStream_t st);   // both passes (two launches)
void grp_next(const GroupLaunch& g, hipStream_t st);
void grp_copy_cols(const GroupLaunch& g, hipStream_t st);

}  // namespace ttx
