// Host side of the ttx engine: workspace management, one decomposition step, the TT and TR sweeps,
// the randomized SVD entry point, synthetic inputs and reconstruction errors.  Everything is queued
// on the handle's stream; the only host synchronisations are the per-step rank decisions of the
// DELTA rank mode and the end of each public call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ttx.h"
#include "ttx_device.h"
#include "ttx_kernels.h"

namespace ttx {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// one grouped launch of the batch engine: kernel kind, template choices, block size and dynamic LDS,
// the job descriptors (bytes) and the block map; joff / moff locate them in the uploaded blob
struct GRec {
  int kind = 0, ar = 0, v0 = 0, v1 = 0, v2 = 0, threads = 256, lds = 0;
  size_t jsize = 0;
  int njobs = 0;
  std::vector<unsigned char> jobs;
  std::vector<int2> map;
  size_t joff = 0, moff = 0;
};
// a batch's grouped plan: the launches in order, the per-tensor sigma / status words, and the graph
// of the launches when the batch runs captured
struct GPlan {
  uint64_t sig = 0, ws = 0;
  std::vector<GRec> recs;
  float* sigstat = nullptr;     // [count][d - 1][TTX_MAX_ELL floats + 4 ints]
  size_t sigstat_bytes = 0;
  int steps = 0, count = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  ~GPlan() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

struct Handle {
  hipStream_t stream = nullptr;
  std::string err;
  DevBuf part, Y, Y2, Qc, Qp, Qn, Z, W1, W2, G, G2, Rinv, lam, V, Ur, gpart, status, next[2], sig, keys, scratch[4];
  DevBuf Yt, tpart, tout;         // held-out test columns of the range check and its partial sums
  // the sketch of the second step of a TT sweep, taken in the first step's Gram pass (the same
  // memory read once): valid for the unfolding (pre_A, pre_m x pre_n) with Omega (seed, stream)
  DevBuf Ypre;
  DevBuf omg;                     // Omega of that pass, precomputed (2 KB chunks of 16 rows, bf16)
  // the current randomized step's sketch with every Omega column (kp of them), reduced and before
  // the fused left factor: Ypre or Y; any basis width takes its first columns (step-local)
  float* Yfull = nullptr;
  // the inverse shifted Cholesky factor (64 x 64) of the Gram of the mixed full sketch of Rprobe_y,
  // taken for the rank model's pivots: its leading w x w block starts the basis of any width w
  DevBuf Rprobe;
  // host-streamed first step: two column-slice buffers, per-slice Gram sums, the remainder
  DevBuf hs[2], hGs, hB, hB2, hA;
  hipStream_t cstream = nullptr;  // host-to-device copies of the slices (default priority)
  // the first step's eigensolve and core beside the second step (a TT sweep whose first pass took
  // both): a stream of their own (default priority), fork / eigenvalues / core events, and buffers
  // the second step does not touch
  hipStream_t estream = nullptr;
  hipEvent_t eev[3] = {nullptr, nullptr, nullptr};
  DevBuf G0, lam0, V0, keys0;
  double* h_lam0 = nullptr;      // pinned, TTX_MAX_ELL
  hipEvent_t hev[4] = {nullptr, nullptr, nullptr, nullptr};   // copied[2], free[2]
  bool rprobe_valid = false;
  const float* rprobe_y = nullptr;
  int rprobe_k = 0;                // the leading rprobe_k x rprobe_k block of Rprobe is computed
  bool pre_valid = false;
  const float* pre_A = nullptr;
  int64_t pre_m = 0, pre_n = 0;
  uint64_t pre_seed = 0;
  uint32_t pre_stream = 0;
  // rank model: parameters, Adam moments, gradient, activations (batch-sized), probe pivots / features
  DevBuf mP, mM1, mM2, mG, mAct, mPart, mIO;
  DevBuf f16s;                    // FP16 operand scales (65 floats) and column-max words (64)
  bool mlp_ready = false;
  int mlp_t = 0;
  int jacobi_max_sweeps = kJacobiMaxSweeps;   // sweep cap of the eigensolver (TTX_JACOBI_MAX_SWEEPS)
  // Omega stream of a sweep's step k is stream_base + k: the host-streamed sweep runs its steps 2..d-1
  // as a sweep of its own with base 1, so they draw the same sketches as the device sweep's
  uint32_t stream_base = 0;
  int* h_status = nullptr;       // pinned
  float* h_sigma = nullptr;      // pinned, TTX_MAX_ELL
  double* h_lam = nullptr;       // pinned
  double* h_aux = nullptr;       // pinned, 16 doubles (range-check sums; words 8.. the rank model output)
  float* h_sweep = nullptr;      // pinned, the sweep's sigma and status words (one copy at the end)
  int dev = 0;
  std::vector<float*> last_cores;  // device buffers kept for the reference-name shims
  // per-kernel timing (TTX_FLAG_TIMING): event pairs recorded on the stream around launches
  bool ktime = false;
  std::vector<hipEvent_t> evpool;
  size_t evnext = 0;
  struct KRec { hipEvent_t a, b; int kind; int64_t bytes; };
  std::vector<KRec> krecs;
  // batches: capture = queue only (no host synchronisation, no readback); strict = under graph
  // capture (no allocation either); sub-handles run the tensors on their own streams
  bool capture = false;
  bool strict = false;
  // batch mode (the workers of a batch, and the plans of the grouped batch engine): the streaming
  // passes take the batch grids (a few large workgroups per tensor, since the whole batch runs at
  // once), sketches take the register-buffered kernel with batch_cpw chunks per workgroup, and a
  // full-rank exact step is projected like any other (no deferred core)
  bool batch_mode = false;
  int64_t batch_cpw = 32;
  std::map<uint64_t, int64_t> cpw_tuned;   // autotuned batch_cpw per batch signature
  // the grouped batch engine: its plan, the tensors' workspaces and the uploaded job blob
  std::unique_ptr<GPlan> gplan;
  DevBuf barena, bblob;
  DevBuf rwork;   // TT rounding: working copies of the cores and a contraction buffer
  std::vector<Handle*> subs;
  hipStream_t bstream = nullptr;             // capture origin of the batch graph (never the null stream)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  uint64_t gsig = 0;
  // workspace generation: bumped whenever ensure() (re)allocates a buffer of this handle, so a graph
  // captured over the workers' buffers is relaunched only while none of them has moved
  uint64_t wsgen = 0;
  uint64_t gws = 0;   // the workers' combined generation when gexec was captured
  // allocation-free mode (ttx_tt_svd_noalloc): ensure() carves buffers from [arena, arena + arena_cap)
  // instead of allocating; noalloc is the parent handle's shadow that runs those calls
  unsigned char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  Handle* noalloc = nullptr;
};

hipError_t noalloc_init(int dev);
void* noalloc_scratch(int dev);

static hipEvent_t next_event(Handle* h) {
  if (h->evnext == h->evpool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    h->evpool.push_back(e);
  }
  return h->evpool[h->evnext++];
}

// Records an event pair around the launches made during its lifetime (TTX_FLAG_TIMING only).
struct KScope {
  Handle* h;
  int kind;
  int64_t bytes;
  hipEvent_t a = nullptr;
  KScope(Handle* h_, int kind_, int64_t bytes_) : h(h_), kind(kind_), bytes(bytes_) {
    if (h->ktime && (a = next_event(h))) (void)hipEventRecord(a, h->stream);
  }
  ~KScope() {
    if (!a) return;
    hipEvent_t b = next_event(h);
    if (!b) return;
    (void)hipEventRecord(b, h->stream);
    h->krecs.push_back({a, b, kind, bytes});
  }
};

static ttx_status fail(Handle* h, ttx_status s, const std::string& msg) {
  if (h) h->err = msg;
  return s;
}

#define TTX_HIP(h, call)                                                                      \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return fail((h), TTX_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

static ttx_status ensure(Handle* h, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return TTX_OK;
  if (h->strict) return fail(h, TTX_ERR_HIP, "workspace growth under graph capture");
  if (h->arena) {
    // a fresh 256-byte aligned piece of the arena; an outgrown piece is not reused within the call
    size_t want = bytes < 256 ? 256 : bytes;
    want = (want + 255) & ~(size_t)255;
    if (want > h->arena_cap - h->arena_used)
      return fail(h, TTX_ERR_UNSUPPORTED, "no-alloc scratch exhausted: " + std::to_string(h->arena_used + want) +
                                              " bytes needed, " + std::to_string(h->arena_cap) + " available");
    b.p = h->arena + h->arena_used;
    b.bytes = want;
    h->arena_used += want;
    return TTX_OK;
  }
  if (b.p) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  ++h->wsgen;
  if (hipMalloc(&b.p, want) != hipSuccess) {
    (void)hipGetLastError();
    return fail(h, TTX_ERR_OOM, "hipMalloc of " + std::to_string(want) + " bytes failed");
  }
  b.bytes = want;
  return TTX_OK;
}

template <typename T>
static T* P(DevBuf& b) { return reinterpret_cast<T*>(b.p); }

static ttx_status check_launch(Handle* h, const char* where) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(h, TTX_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
  return TTX_OK;
}

#define TTX_TRY(x)                 \
  do {                             \
    ttx_status s_ = (x);           \
    if (s_ != TTX_OK) return s_;   \
  } while (0)

static int kp_for(int k) { return k <= 32 ? 32 : 64; }

// Grids of a batch's streaming passes.  The whole batch runs at once, so each tensor takes a few large
// workgroups instead of filling the machine by itself: the exact Gram 8192 columns per workgroup, the
// projection about 2 MB of A per workgroup (2048 / m column blocks of 256).
static int64_t gram_small_grid_batch(int64_t n) {
  const int64_t g = (n + 8191) / 8192;
  return g < 1 ? 1 : g;
}
static int64_t proj_grid_batch(int64_t m, int64_t n) {
  const int64_t bpw = m >= 2048 ? 1 : 2048 / m;
  const int64_t g = (proj_blocks(n) + bpw - 1) / bpw;
  return g < 1 ? 1 : g;
}
static int64_t gram_grid_for(const Handle* h, int64_t n) { return h->batch_mode ? gram_small_grid_batch(n) : gram_small_grid(n); }
static int64_t proj_grid_for(const Handle* h, int64_t m, int64_t n) { return h->batch_mode ? proj_grid_batch(m, n) : proj_grid(n); }

// ------------------------------------------------------------------------------------------------
// arithmetic choices
// ------------------------------------------------------------------------------------------------
struct Arith3 {
  int range;    // sketch and power passes
  int accurate; // remainder / exact Gram / exact projection
  int tail;     // the randomized step's final projection Z = A^T Q
};
static Arith3 arith_for(int precision, int flags) {
  Arith3 a;
  if (precision == TTX_PREC_FP32) { a.range = AR_FP32; a.accurate = AR_FP32; }
  else if (precision == TTX_PREC_FP16) { a.range = AR_FP16; a.accurate = AR_BF16X4; }
  else { a.range = (flags & TTX_FLAG_FAST_RANGE) ? AR_BF16 : AR_BF16X3; a.accurate = AR_BF16X4; }
  a.tail = (precision == TTX_PREC_BF16 && (flags & TTX_FLAG_TAIL_X3)) ? AR_BF16X3 : a.accurate;
  return a;
}

// ------------------------------------------------------------------------------------------------
// orthonormal basis of the columns of Y (m x k, row-major stride kpy): shifted CholeskyQR then two
// CholeskyQR rounds, all in FP64; output Qc column-major (kp x m), columns k..kp-1 zero
// ------------------------------------------------------------------------------------------------
static ttx_status orth(Handle* h, const float* Y, int64_t kpy, int m, int k, int kp, float* Qc, int* stat,
                       const double* Rlead = nullptr, int ldr = 0) {
  hipStream_t st = h->stream;
  const int64_t nb = panel_blocks(m);
  TTX_TRY(ensure(h, h->gpart, (size_t)nb * k * k * sizeof(double)));
  TTX_TRY(ensure(h, h->Rinv, (size_t)k * k * sizeof(double)));
  TTX_TRY(ensure(h, h->W1, (size_t)m * k * sizeof(double)));
  TTX_TRY(ensure(h, h->W2, (size_t)m * k * sizeof(double)));
  double* Ri = P<double>(h->Rinv);
  double* W1 = P<double>(h->W1);
  double* W2 = P<double>(h->W2);
  double* gp = P<double>(h->gpart);
  KScope ks(h, TTX_K_SMALL, 0);
  if (Rlead) {
    // the first (shifted) round's factor is the leading k x k block of a factor already taken
    launch_apply_gram_f32(Y, kpy, m, k, Rlead, W1, gp, st, ldr);
  } else {
    launch_gram_tall_f32(Y, kpy, m, k, gp, nb, st);
    launch_chol_inv(gp, nb, k, m, 1, Ri, stat, st);
    launch_apply_gram_f32(Y, kpy, m, k, Ri, W1, gp, st);
  }
  launch_chol_inv(gp, nb, k, m, 0, Ri, stat, st);
  launch_apply_gram_f64(W1, m, k, Ri, W2, gp, st);
  // after two CholeskyQR rounds ||W2^T W2 - I|| is at rounding level: one Newton-Schulz step
  // (polar factor, same column span) replaces a third Cholesky
  launch_apply_ns_out(W2, m, k, kp, gp, nb, Qc, m, st);
  return check_launch(h, "orth");
}

// ------------------------------------------------------------------------------------------------
// one step
// ------------------------------------------------------------------------------------------------
struct StepArgs {
  const float* A = nullptr;
  int64_t m = 0, n = 0;
  int r = 1;            // rank (fixed) or ceiling
  int s = 10, q = 2;
  uint64_t seed = 1234;
  uint32_t stream_id = 0;
  int precision = TTX_PREC_BF16;
  int method = TTX_METHOD_AUTO;
  int flags = 0;
  int rank_mode = TTX_RANK_FIXED;
  double delta2 = 0.0;  // absolute squared truncation threshold for DELTA (< 0: delta_rel times this
                        // step's total, known only after its exact Gram)
  double delta_rel = 0.0;   // eps^2 / (d - 1)
  bool allow_defer = false; // DELTA / LEARNED: leave a full-rank exact step unprojected when r = m
  double norm2 = -1.0;  // ||A||_F^2 when known
  float* core = nullptr;   // m x r_out
  float* next = nullptr;   // r_out x n
  // fused full-rank step: A = (I_N (x) pend^T) S where S = this->A and pend is the previous step's
  // square core (pend_ms x pend_ms); defer: compute the core only and leave A_next implicit
  const float* pend = nullptr;
  int pend_ms = 0;
  bool defer = false;
  // device slots for sigma (TTX_MAX_ELL floats) and status (4 ints) of this step; sync: copy them to
  // the host at the end of the step (single-step entry points)
  float* sig_dev = nullptr;
  int* stat_dev = nullptr;
  bool sync = true;
  double rel_budget = 0.0;   // LEARNED: delta^2 / ||A_k||^2 (a model input)
  // previous step's kept singular values on the device (FP16 operand scale of this step's input)
  const float* sig_prev = nullptr;
  int r_prev = 0;
  // TTX_FLAG_ADAPTIVE: relative squared tolerance of the range check; pred_ell > 0: the sketch width
  // a learned model predicted (the check judges it, the adaptive growth is its fallback)
  double range_tol2 = 0.0;
  int pred_ell = 0;
  // exact first step whose Gram partials (gram_pre of them, in gpart) the fused first pass made
  int64_t gram_pre = 0;
  // TR first step (exact path): tr_r1c > 0 writes the remainder straight into the 231 layout, rows
  // c = a + r0 b with b < tr_r1c only (the rows beyond the first step's rank bound are zero)
  int tr_r0 = 0, tr_r1c = 0;
  // rows the path decision is made on (0: m): the compacted second TR step decides as its zero-padded
  // unfolding would
  int64_t m_path = 0;
  // fused full-rank step whose pending square core is applied late: the step runs on S itself.  With
  // P = pend_core orthogonal (pend_ms x pend_ms), A = (I_N (x) P^T) S and S have the same sketch Gram,
  // range residual, Z = A^T Q_A = S^T Q_S, eigenvectors and remainder, Q_A = (I_N (x) P^T) Q_S; only
  // the core G = Q_A U_B needs P, so join() runs before the rank decision and the assembly: it waits
  // for P, returns this step's norm2 / delta2 (known once the previous step's eigenvalues are), and
  // returns TTX_RETRY when the previous step kept fewer than pend_ms singular values after all
  const float* pend_core = nullptr;
  std::function<ttx_status(double*, double*)> join;
  // DELTA / LEARNED ranks: the chosen rank is rounded up to a multiple of this (the TR first step keeps
  // r0 * r1 columns: r0 * ceil(R / r0) for a delta rank R)
  int rank_multiple = 1;
  // exact paths only: the eigensolver's Gram (A A^T of the exact-m path, Z^T Z of the tail) summed in
  // FP64 from the FP32 operands instead of FP32 MFMA products, so singular values far below the largest
  // keep their accuracy (sqrt of the Gram's relative error: about 1e-8 instead of 3e-4 of sigma_1)
  bool precise = false;
};
// internal only, never returned to a caller: the step ran on a speculated left factor that did not
// come true and is to be run again (7: above every public code, inside the enumeration's range)
static constexpr ttx_status TTX_RETRY = static_cast<ttx_status>(7);
static_assert(TTX_ERR_UNSUPPORTED < 7, "TTX_RETRY collides with a public status");
struct StepOut {
  int r = 0, ell = 0, path = 0, power_iters = 0, zero_pivots = 0, sweeps = 0, kdim = 0;
  int nonconv = 0;         // the eigensolver reached its sweep cap before converging
  double kept2 = 0.0;      // sum of kept sigma^2 = ||A_next||_F^2
  double total2 = -1.0;    // ||A||_F^2 when the step knows it
  int64_t bytes = 0;
  double resid = -1.0;     // relative range residual estimate of the basis used (-1: not checked)
  int checks = 0, passes = 0, fallback = 0, pred_rank = 0;
  bool deferred = false;   // the step left A_next implicit (full-rank exact step)
  int tr_r1c = 0;          // TR first step: rows b of the 231 layout its projection wrote
  float sigma[TTX_MAX_ELL];
};

// host copies of a step's sigma (kdim values) and status into its StepOut
// singular values and solver status of a finished step; with norms = true also the kept and total
// squared norms from the FP32 singular values (a DELTA / LEARNED step sets those from its FP64
// eigenvalues when it chooses the rank)
static void finish_step_out(const float* sig, const int* stat, StepOut& o, bool norms = true) {
  double total = 0.0, kept = 0.0;
  for (int i = 0; i < std::min(o.kdim, TTX_MAX_ELL); ++i) {
    o.sigma[i] = sig[i];
    const double s2 = (double)sig[i] * sig[i];
    total += s2;
    if (i < o.r) kept += s2;
  }
  if (norms) {
    o.kept2 = kept;
    if (o.path == TTX_PATH_EXACT_M) o.total2 = total;
  }
  o.zero_pivots = stat[0];
  o.sweeps = stat[1];
  o.nonconv = stat[2];
}

// rank of the delta rule: the smallest r with tail(r) <= delta^2 (Oseledets 2011, TT-SVD; the
// commented rule of trdec_gpu_float_unopt2_rank.cu:222-243).  rel2 < 0: tail(r) = max(total2 - sum_{i<=r}
// lam_i, sum_{i>r} lam_i) with total2 = ||A||_F^2 carried from the previous step.  rel2 >= 0 (the
// basis' relative range residual, 0 for a complete basis): tail(r) = sum_{i>r} lam_i + rel2 / (1 - rel2)
// sum_i lam_i, which needs no cancellation against a norm carried across steps (whose rounding in
// reduced-precision passes can exceed delta^2).
static int choose_rank(const StepArgs& a, int k, const double* lam, double total2, double rel2, double* kept2) {
  int r = a.r;
  if (a.rank_mode != TTX_RANK_FIXED) {
    double all = 0.0;
    for (int i = 0; i < k; ++i) all += std::max(lam[i], 0.0);
    const double outside = rel2 >= 0.0 ? (rel2 < 1.0 ? rel2 / (1.0 - rel2) * all : 1e300) : 0.0;
    double cum = 0.0;
    r = k;
    for (int i = 0; i < k; ++i) {
      cum += std::max(lam[i], 0.0);
      const double cap_tail = all - cum;
      const double tail = rel2 >= 0.0 ? cap_tail + outside : std::max(total2 - cum, cap_tail);
      if (tail <= a.delta2) { r = i + 1; break; }
    }
    r = std::min(r, a.r);
    r = std::max(r, 1);
    if (a.rank_multiple > 1) r = std::min(a.r, (r + a.rank_multiple - 1) / a.rank_multiple * a.rank_multiple);
  }
  double s = 0.0;
  for (int i = 0; i < std::min(r, k); ++i) s += std::max(lam[i], 0.0);
  *kept2 = s;
  return r;
}

// operand of the projections: Q itself, or P^T Q for a fused step
static const float* proj_operand(Handle* h, const StepArgs& a, const float* Qsrc, int kp) {
  if (!a.pend) return Qsrc;
  {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_blockmix(a.pend, a.pend_ms, (int)a.m, kp, 1, Qsrc, P<float>(h->Qp), h->stream);
  }
  return P<float>(h->Qp);
}

// tail of a randomized / exact-n step: Z = A^T Q (accurate), eig of Z^T Z, core and remainder
static ttx_status step_tail(Handle* h, const StepArgs& a, const Arith3& ar, int k, int kp, double rel2,
                            StepOut& o) {
  hipStream_t st = h->stream;
  const int m = (int)a.m;
  const int64_t n = a.n;
  TTX_TRY(ensure(h, h->Z, (size_t)n * kp * sizeof(float)));
  float* Z = P<float>(h->Z);
  const int kz = (k + 3) & ~3;   // Z keeps only the basis columns (rounded to 16 bytes)
  // Z = A^T Q (FP32-accurate) and its Gram B B^T = Z^T Z in the same pass
  int64_t gg = proj_grid_for(h, m, n);
  TTX_TRY(ensure(h, h->gpart, (size_t)gg * k * k * sizeof(double)));
  const float* Qop = proj_operand(h, a, P<float>(h->Qc), kp);
  {
    KScope ks(h, TTX_K_PROJ, a.m * n * 4 + n * kz * 4);
    launch_proj_grid(ar.tail, kp, a.A, a.m, m, n, Qop, m, 0, Z, kz, kz, gg, st, a.precise ? nullptr : P<double>(h->gpart),
                     a.precise ? 0 : k);
  }
  if (a.precise) {
    // Z^T Z in FP64 from the FP32 rows of Z
    gg = gram_tall_grid(n);
    TTX_TRY(ensure(h, h->gpart, (size_t)gg * k * k * sizeof(double)));
    KScope ks(h, TTX_K_SMALL, 0);
    launch_gram_tall_f32(Z, kz, n, k, P<double>(h->gpart), gg, st);
  }
  o.bytes += (int64_t)(a.m * n * 4);
  o.passes += 1;
  {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_reduce_f64(P<double>(h->gpart), gg, (int64_t)k * k, P<double>(h->G), st);
    launch_syevj(P<double>(h->G), k, P<double>(h->lam), P<double>(h->V), a.stat_dev,
                 P<unsigned long long>(h->keys), st, h->jacobi_max_sweeps);
  }
  TTX_TRY(check_launch(h, "step tail"));
  StepArgs aj = a;
  if (a.pend_core && a.join) TTX_TRY(a.join(&aj.norm2, &aj.delta2));
  // rank decision on the host when needed
  int r = a.r;
  if (a.rank_mode != TTX_RANK_FIXED) {
    TTX_HIP(h, hipMemcpyAsync(h->h_lam, h->lam.p, k * sizeof(double), hipMemcpyDeviceToHost, st));
    TTX_HIP(h, hipStreamSynchronize(st));
    double kept;
    r = choose_rank(aj, k, h->h_lam, aj.norm2, rel2, &kept);
    o.kept2 = kept;
  }
  if (r > k) return fail(h, TTX_ERR_BAD_ARG, "rank exceeds the sketch width");
  TTX_TRY(ensure(h, h->Ur, (size_t)((k + 3) & ~3) * next_from_z_ld(r) * sizeof(float)));
  // G = Q U_B[:, :r] with the sign gauge; sigma = sqrt(eig(B B^T)) (basicRandSVD_LU :717-763); a late
  // left factor turns Q_S into Q_A first (the gauge is taken on the core itself)
  const float* Qasm = P<float>(h->Qc);
  if (a.pend_core) {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_blockmix(a.pend_core, a.pend_ms, m, kp, 2, P<float>(h->Qc), P<float>(h->Qp), st);
    Qasm = P<float>(h->Qp);
  }
  {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_assemble(Qasm, m, m, k, P<double>(h->lam), P<double>(h->V), r, a.core, P<float>(h->Ur),
                    nullptr, 0, a.sig_dev, P<unsigned long long>(h->keys), st);
  }
  // remainder S V^T = U_r^T B = U_r^T Z^T (:734-739)
  {
    KScope ks(h, TTX_K_TALL, n * kz * 4 + (int64_t)r * n * 4);
    launch_next_from_z(Z, kz, n, k, P<float>(h->Ur), r, a.next, st, h->batch_mode);
  }
  o.bytes += (int64_t)(n * kz * 4 + (int64_t)r * n * 4);
  o.r = r;
  return check_launch(h, "step tail assemble");
}

// Y of a fused step: Y_A = P Y_S (in Y2); otherwise Y itself
static float* left_mix(Handle* h, const StepArgs& a, float* Y, int kp) {
  if (!a.pend) return Y;
  {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_blockmix(a.pend, a.pend_ms, (int)a.m, kp, 0, Y, P<float>(h->Y2), h->stream);
  }
  return P<float>(h->Y2);
}

// the sketch's tiling: a batch's plan in batch mode, else the whole-machine plan
static ttx_status plan_for(Handle* h, int ar, const float* A, int m, int64_t n, int kp, SketchPlan* out) {
  (void)ar;
  (void)A;
  *out = h->batch_mode ? plan_sketch_batch(m, n, kp, h->batch_cpw) : plan_sketch(m, n, kp, 0);
  return TTX_OK;
}

// FP16 range finder: power-of-two operand scales (A from the previous step's kept energy)
static float* fp16_scales(Handle* h, const StepArgs& a, const Arith3& ar, bool set_a) {
  if (ar.range != AR_FP16) return nullptr;
  if (!h->f16s.p) {
    if (ensure(h, h->f16s, 129 * sizeof(float)) != TTX_OK) return nullptr;
    launch_zero(h->f16s.p, 129 * sizeof(float), h->stream);
  }
  float* sc = P<float>(h->f16s);
  if (set_a) launch_fp16_ascale(a.sig_prev, a.r_prev, (double)a.m * (double)a.n, sc, h->stream);
  return sc;
}

// Y = A Omega[:, :w + pt] (basicRandSVD_LU :585-620; Omega from the project's Philox stream), test
// columns w .. w+pt-1 kept in Yt, Q = orth(Y[:, :w])
static ttx_status range_sketch(Handle* h, const StepArgs& a, const Arith3& ar, int w, int pt, int kp,
                               const SketchPlan& sp, StepOut& o) {
  hipStream_t st = h->stream;
  const int m = (int)a.m;
  float* Y = P<float>(h->Y);
  if (h->Yfull) {
    // the step's sketch with every Omega column: its first w + pt columns are this sketch (column c
    // of Omega does not depend on the width)
    Y = h->Yfull;
  } else {
    float* sc = fp16_scales(h, a, ar, true);
    {
      KScope ks(h, TTX_K_SKETCH, a.m * a.n * 4);
      launch_sketch(ar.range, sp, a.A, a.m, m, a.n, nullptr, 0, a.seed, a.stream_id, w + pt, P<float>(h->part), st,
                    sc);
    }
    {
      KScope ks(h, TTX_K_SMALL, 0);
      launch_reduce_f32(P<float>(h->part), sp.slabs, (int64_t)m * kp, Y, st);
    }
    TTX_TRY(check_launch(h, "sketch"));
    o.bytes += a.m * a.n * 4;
    o.passes += 1;
  }
  float* Ya = left_mix(h, a, Y, kp);
  if (pt) {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_copy_block(Ya, kp, m, w, pt, P<float>(h->Yt), st);
  }
  // the rank model's factor of this same mixed sketch (all kp = 64 columns, shift for 64 columns)
  // gives the first round for any width: the Cholesky factor of a leading principal block is the
  // leading block of the factor, and so is its inverse
  const bool reuse = h->rprobe_valid && Y == h->rprobe_y && kp == 64 && w <= h->rprobe_k;
  return orth(h, Ya, kp, m, w, kp, P<float>(h->Qc), a.stat_dev, reuse ? P<double>(h->Rprobe) : nullptr, 64);
}

// one power iteration on the current basis: Z = A^T Q, Y = A Z, Q = orth(Y)
static ttx_status power_iter(Handle* h, const StepArgs& a, const Arith3& ar, int k, int kp, const SketchPlan& sp,
                             StepOut& o) {
  hipStream_t st = h->stream;
  const int m = (int)a.m;
  const int64_t n = a.n;
  float* Y = P<float>(h->Y);
  const float* Qop = proj_operand(h, a, P<float>(h->Qc), kp);
  float* sc = fp16_scales(h, a, ar, false);
  {
    KScope ks(h, TTX_K_PROJ, a.m * n * 4 + n * kp * 4);
    // FP16: the projection also takes Z's column maxima (the power pass's operand scales)
    launch_proj_grid(ar.range, kp, a.A, a.m, m, n, Qop, m, 0, P<float>(h->Z), kp, 0, proj_grid_for(h, m, n), st,
                     nullptr, 0, sc, sc ? reinterpret_cast<unsigned*>(sc + 65) : nullptr);
  }
  if (sc) {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_fp16_colscale(reinterpret_cast<unsigned*>(sc + 65), kp, sc, st);
  }
  {
    KScope ks(h, TTX_K_SKETCH, a.m * n * 4 + n * kp * 4);
    launch_sketch(ar.range, sp, a.A, a.m, m, n, P<float>(h->Z), kp, 0, 0, k, P<float>(h->part), st, sc);
  }
  {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_reduce_f32(P<float>(h->part), sp.slabs, (int64_t)m * kp, Y, st);
  }
  TTX_TRY(check_launch(h, "power iteration"));
  o.bytes += 2 * a.m * n * 4;
  o.passes += 2;
  return orth(h, left_mix(h, a, Y, kp), kp, m, k, kp, P<float>(h->Qc), a.stat_dev);
}

// relative squared range residual estimate sum ||(I - Q Q^T) Y_t||^2 / sum ||Y_t||^2 (host sync)
static ttx_status range_check(Handle* h, const StepArgs& a, int k, int pt, double* rel2, StepOut& o) {
  hipStream_t st = h->stream;
  const int m = (int)a.m;
  const int64_t nb = panel_blocks(m);
  TTX_TRY(ensure(h, h->tpart, (size_t)nb * k * pt * sizeof(double)));
  TTX_TRY(ensure(h, h->tout, (size_t)(2 * nb + 2) * sizeof(double)));
  double* tout = P<double>(h->tout);
  {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_test_resid(P<float>(h->Qc), m, m, k, P<float>(h->Yt), pt, P<double>(h->tpart), tout, st);
    launch_reduce_f64(tout, nb, 2, tout + 2 * nb, st);
  }
  TTX_TRY(check_launch(h, "range check"));
  TTX_HIP(h, hipMemcpyAsync(h->h_aux, tout + 2 * nb, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  TTX_HIP(h, hipStreamSynchronize(st));
  const double r2 = h->h_aux[0], y2 = h->h_aux[1];
  *rel2 = y2 > 0.0 ? r2 / y2 : 0.0;
  o.checks += 1;
  return TTX_OK;
}

// ------------------------------------------------------------------------------------------------
// rank model
// ------------------------------------------------------------------------------------------------
static constexpr int MLP_IN = TTX_MLP_IN, MLP_H = TTX_MLP_HIDDEN, MLP_OUT = TTX_MLP_OUT;
static constexpr int PROBE_K = 64;

// activation buffer for batch B: Z1, H1, Z2, H2, dH, dZ (B x H each), Y, dY (B x OUT each)
static ttx_status mlp_buffers(Handle* h, int B) {
  TTX_TRY(ensure(h, h->mP, TTX_MLP_PARAMS * sizeof(float)));
  TTX_TRY(ensure(h, h->mM1, TTX_MLP_PARAMS * sizeof(float)));
  TTX_TRY(ensure(h, h->mM2, TTX_MLP_PARAMS * sizeof(float)));
  TTX_TRY(ensure(h, h->mG, TTX_MLP_PARAMS * sizeof(float)));
  TTX_TRY(ensure(h, h->mAct, ((size_t)6 * B * MLP_H + (size_t)2 * B * MLP_OUT) * sizeof(float)));
  TTX_TRY(ensure(h, h->mPart, (size_t)(mlp_loss_blocks((int64_t)B * MLP_OUT) + 1) * sizeof(double)));
  TTX_TRY(ensure(h, h->mIO, (PROBE_K * sizeof(double) + (MLP_IN + MLP_OUT) * sizeof(float))));
  return TTX_OK;
}

struct MlpAct {
  float *Z1, *H1, *Z2, *H2, *dH, *dZ, *Y, *dY;
};
static MlpAct mlp_act(Handle* h, int B) {
  float* b = P<float>(h->mAct);
  const size_t bh = (size_t)B * MLP_H, bo = (size_t)B * MLP_OUT;
  MlpAct a;
  a.Z1 = b; a.H1 = b + bh; a.Z2 = b + 2 * bh; a.H2 = b + 3 * bh; a.dH = b + 4 * bh; a.dZ = b + 5 * bh;
  a.Y = b + 6 * bh; a.dY = a.Y + bo;
  return a;
}

// the probe: pivots of the shifted Cholesky of the Gram of Y_p = A[:, :n_p] Omega (64 columns)
// Ysrc != nullptr: the pivots of that sketch (the step's own 64-column sketch of every column of A,
// reduced, before the fused left factor) instead of a probe pass
static ttx_status probe(Handle* h, const StepArgs& a, const Arith3& ar, double* d_piv, StepOut& o,
                        const float* Ysrc = nullptr, int kinv = 0) {
  hipStream_t st = h->stream;
  const int m = (int)a.m;
  if (Ysrc) {
    const int kp = PROBE_K;
    const int64_t nb = panel_blocks(m);
    TTX_TRY(ensure(h, h->gpart, (size_t)nb * kp * kp * sizeof(double)));
    float* Ya = left_mix(h, a, const_cast<float*>(Ysrc), kp);
    TTX_TRY(ensure(h, h->Rprobe, (size_t)kp * kp * sizeof(double)));
    KScope ks(h, TTX_K_SMALL, 0);
    launch_gram_tall_f32(Ya, kp, m, kp, P<double>(h->gpart), nb, st);
    // with the inverse: the basis of the chosen width starts from its leading block
    launch_chol_inv(P<double>(h->gpart), nb, kp, m, 1, P<double>(h->Rprobe), nullptr, st, d_piv, kinv);
    h->rprobe_valid = true;
    h->rprobe_y = Ysrc;
    h->rprobe_k = kinv > 0 ? kinv : kp;
    return check_launch(h, "sketch pivots");
  }
  int64_t np = std::max<int64_t>(std::min<int64_t>(a.n, 512), a.n / 16);
  if (np > a.n) np = a.n;
  const int kp = PROBE_K;
  const SketchPlan sp = plan_sketch(m, np, kp, 0);
  TTX_TRY(ensure(h, h->part, (size_t)sp.slabs * m * kp * sizeof(float)));
  TTX_TRY(ensure(h, h->Y, (size_t)m * kp * sizeof(float)));
  TTX_TRY(ensure(h, h->Y2, (size_t)m * kp * sizeof(float)));
  // FP32-accurate against the exact Omega in every precision mode
  const int range = (ar.range == AR_FP32) ? AR_FP32 : AR_BF16X3;
  {
    KScope ks(h, TTX_K_SKETCH, a.m * np * 4);
    launch_sketch(range, sp, a.A, a.m, m, np, nullptr, 0, a.seed, a.stream_id | 0x10000u, kp, P<float>(h->part), st);
  }
  const int64_t nb = panel_blocks(m);
  TTX_TRY(ensure(h, h->gpart, (size_t)nb * kp * kp * sizeof(double)));
  {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_reduce_f32(P<float>(h->part), sp.slabs, (int64_t)m * kp, P<float>(h->Y), st);
  }
  float* Ya = left_mix(h, a, P<float>(h->Y), kp);
  {
    KScope ks(h, TTX_K_SMALL, 0);
    launch_gram_tall_f32(Ya, kp, m, kp, P<double>(h->gpart), nb, st);
    launch_chol_inv(P<double>(h->gpart), nb, kp, m, 1, nullptr, nullptr, st, d_piv);
  }
  o.bytes += a.m * np * 4;
  return check_launch(h, "probe");
}

// predicted (rank, width) of one randomized step; host sync
static ttx_status predict(Handle* h, const StepArgs& a, const Arith3& ar, int ell, int* pred_r, int* pred_w,
                          StepOut& o, const float* Ysrc = nullptr) {
  if (!h->mlp_ready) return fail(h, TTX_ERR_BAD_ARG, "rank mode LEARNED needs a rank model (ttx_mlp_load)");
  hipStream_t st = h->stream;
  TTX_TRY(mlp_buffers(h, 1));
  double* piv = P<double>(h->mIO);
  float* x = reinterpret_cast<float*>(piv + PROBE_K);
  float* y = x + MLP_IN;
  // the basis widths tried afterwards are at most ell: only that leading block of the factor's
  // inverse is needed
  TTX_TRY(probe(h, a, ar, piv, o, Ysrc, ell));
  {
    KScope ks(h, TTX_K_SMALL, 0);
    mlp_predict(P<float>(h->mP), piv, PROBE_K, a.rel_budget > 0 ? a.rel_budget : 1e-12, (double)a.m, (double)a.n, y,
                st);
  }
  TTX_TRY(check_launch(h, "rank model"));
  // through the pinned scratch words (a pageable destination would add a staging copy)
  float* hy = reinterpret_cast<float*>(h->h_aux + 8);
  static_assert(MLP_OUT * sizeof(float) <= 8 * sizeof(double), "rank model output fits the pinned words");
  TTX_HIP(h, hipMemcpyAsync(hy, y, MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, st));
  TTX_HIP(h, hipStreamSynchronize(st));
  const double r = std::isfinite(hy[0]) ? 64.0 * hy[0] : 1.0;
  const double w = std::isfinite(hy[1]) ? 64.0 * hy[1] : (double)ell;
  *pred_r = std::max(1, std::min(a.r, (int)std::lround(r)));
  // four columns of margin over the predicted width: on 32 order-5 tensors (64 steps) the range check
  // failed for 36% of the steps with two, 9% with four and 2% with six columns, and four cost no time
  // against two on either the learned or the 32^6 workload (six: 2% slower on the learned one)
  *pred_w = std::max(1, std::min(ell, (int)std::ceil(w) + 4));
  return TTX_OK;
}

static ttx_status run_step(Handle* h, const StepArgs& a_in, StepOut& o) {
  hipStream_t st = h->stream;
  StepArgs a = a_in;
  const Arith3 ar = arith_for(a.precision, a.flags);
  const int m = (int)a.m;
  const int64_t n = a.n;
  if (a.m <= 0 || a.n <= 0 || a.r <= 0) return fail(h, TTX_ERR_BAD_ARG, "step: empty unfolding or rank");
  if (a.m > (int64_t)1 << 30) return fail(h, TTX_ERR_UNSUPPORTED, "step: more than 2^30 rows");
  if ((a.pend || a.pend_core) && (a.pend_ms <= 0 || a.m % a.pend_ms != 0 || a.pend_ms > TTX_MAX_ELL))
    return fail(h, TTX_ERR_BAD_ARG, "step: bad fused left factor");
  if (a.pend_core) a.pend = nullptr;   // the step runs on S; the factor reaches the core only
  int ell = a.r + a.s;
  const int64_t mpath = a.m_path > 0 ? a.m_path : a.m;
  const int64_t mn = std::min<int64_t>(mpath, a.n);
  // AUTO: exact when the sketch would be as wide as the short side; in batch mode also whenever the
  // short side is at most TTX_MAX_ELL (one streaming pass for its Gram, or none for its basis, against
  // 2 q + 2 for the randomized range finder)
  const bool exact = (a.method == TTX_METHOD_EXACT) ||
                     (a.method == TTX_METHOD_AUTO && (ell >= mn || (h->batch_mode && mn <= TTX_MAX_ELL)));
  memset(o.sigma, 0, sizeof(o.sigma));
  o.bytes = 0;
  TTX_TRY(ensure(h, h->status, 4 * sizeof(int)));
  TTX_TRY(ensure(h, h->sig, TTX_MAX_ELL * sizeof(float)));
  if (!a.stat_dev) {
    a.stat_dev = P<int>(h->status);
    TTX_HIP(h, hipMemsetAsync(h->status.p, 0, 4 * sizeof(int), st));
  }
  if (!a.sig_dev) a.sig_dev = P<float>(h->sig);
  TTX_TRY(ensure(h, h->lam, TTX_MAX_ELL * sizeof(double)));
  TTX_TRY(ensure(h, h->keys, 64 * sizeof(unsigned long long)));
  TTX_TRY(ensure(h, h->V, TTX_MAX_ELL * TTX_MAX_ELL * sizeof(double)));
  TTX_TRY(ensure(h, h->G, TTX_MAX_ELL * TTX_MAX_ELL * sizeof(double)));
  TTX_TRY(ensure(h, h->G2, TTX_MAX_ELL * TTX_MAX_ELL * sizeof(double)));
  int k = 0;
  if (exact && mpath <= TTX_MAX_ELL) {
    // exact Gram path on the short side (ttdec6_singleGPU.cu:510-536): C = A A^T, eig, top r, G^T A
    if (a.pend_core) {
      // this path mixes the Gram itself: the left factor is needed up front
      if (a.join) TTX_TRY(a.join(&a.norm2, &a.delta2));
      a.pend = a.pend_core;
      a.pend_core = nullptr;
      a.join = nullptr;
    }
    o.path = TTX_PATH_EXACT_M;
    o.ell = 0;
    k = m;
    int64_t grid = gram_grid_for(h, n);
    if (a.gram_pre > 0) {
      grid = a.gram_pre;   // the fused first pass left them in gpart
    } else if (a.precise) {
      // A A^T as the FP64 Gram of the rows of A^T (A column-major m x n is A^T row-major, stride m)
      grid = gram_tall_grid(n);
      TTX_TRY(ensure(h, h->gpart, (size_t)grid * m * m * sizeof(double)));
      KScope ks(h, TTX_K_GRAM, a.m * n * 4);
      launch_gram_tall_f32(a.A, m, n, m, P<double>(h->gpart), grid, st);
    } else {
      TTX_TRY(ensure(h, h->gpart, (size_t)grid * m * m * sizeof(double)));
      KScope ks(h, TTX_K_GRAM, a.m * n * 4);
      launch_gram_small(ar.accurate, a.A, a.m, m, n, P<double>(h->gpart), grid, nullptr, st);
    }
    {
      KScope ks(h, TTX_K_SMALL, 0);
      if (a.pend) {
        launch_reduce_f64(P<double>(h->gpart), grid, (int64_t)m * m, P<double>(h->G2), st);
        launch_gram_mix(a.pend, a.pend_ms, m, P<double>(h->G2), P<double>(h->G), st);
      } else {
        launch_reduce_f64(P<double>(h->gpart), grid, (int64_t)m * m, P<double>(h->G), st);
      }
      launch_syevj(P<double>(h->G), m, P<double>(h->lam), P<double>(h->V), a.stat_dev,
                   P<unsigned long long>(h->keys), st, h->jacobi_max_sweeps);
    }
    TTX_TRY(check_launch(h, "exact gram"));
    o.bytes += a.m * n * 4;
    o.passes += 1;
    int r = a.r;
    if (a.rank_mode != TTX_RANK_FIXED) {
      TTX_HIP(h, hipMemcpyAsync(h->h_lam, h->lam.p, m * sizeof(double), hipMemcpyDeviceToHost, st));
      TTX_HIP(h, hipStreamSynchronize(st));
      double total = 0.0;
      for (int i = 0; i < m; ++i) total += std::max(h->h_lam[i], 0.0);
      StepArgs aa = a;
      if (aa.norm2 < 0) aa.norm2 = total;
      if (aa.delta2 < 0) aa.delta2 = a.delta_rel * total;
      double kept = 0.0;
      r = choose_rank(aa, m, h->h_lam, total, 0.0, &kept);
      o.total2 = total;
      o.kept2 = kept;
    }
    if (r > TTX_MAX_ELL) return fail(h, TTX_ERR_UNSUPPORTED, "rank above TTX_MAX_ELL");
    const int kpn = kp_for(r);
    TTX_TRY(ensure(h, h->Qn, (size_t)kpn * m * sizeof(float)));
    TTX_TRY(ensure(h, h->Qp, (size_t)kpn * m * sizeof(float)));
    {
      KScope ks(h, TTX_K_SMALL, 0);
      launch_assemble(nullptr, 0, m, m, P<double>(h->lam), P<double>(h->V), r, a.core, nullptr, P<float>(h->Qn),
                      kpn, a.sig_dev, P<unsigned long long>(h->keys), st);
    }
    o.deferred = a.defer || (a.allow_defer && r == m);
    if (!o.deferred) {
      const float* Qop = proj_operand(h, a, P<float>(h->Qn), kpn);
      // rows written: r, or (TR first step, relaid out on the way) r0 times the kept b range (the
      // rows b < r / r0 that the rank keeps, and of those the ones that can be nonzero)
      const int r1c = a.tr_r1c > 0 ? std::min(a.tr_r1c, r / a.tr_r0) : 0;
      o.tr_r1c = r1c;
      const int64_t rw = r1c > 0 ? (int64_t)a.tr_r0 * r1c : r;
      KScope ks(h, TTX_K_PROJ, a.m * n * 4 + rw * n * 4);
      // the exact path's m <= 64 rows: the short-unfolding projection
      if (r1c > 0) launch_proj_short(ar.accurate, kpn, a.A, a.m, m, n, Qop, m, 2, a.next, a.tr_r0, r1c, 0, st);
      else launch_proj_short(ar.accurate, kpn, a.A, a.m, m, n, Qop, m, 1, a.next, 0, r,
                             h->batch_mode ? proj_short_grid(n, true) : 0, st);
      o.bytes += a.m * n * 4 + rw * n * 4;
      o.passes += 1;
    }
    o.r = r;
    TTX_TRY(check_launch(h, "exact projection"));
  } else if (exact && a.n <= TTX_MAX_ELL) {
    // short-and-wide on the other side: Q = orth(A), then the common tail
    o.path = TTX_PATH_EXACT_N;
    o.ell = 0;
    k = (int)n;
    const int kp = kp_for(k);
    TTX_TRY(ensure(h, h->Y, (size_t)m * kp * sizeof(float)));
    TTX_TRY(ensure(h, h->Y2, (size_t)m * kp * sizeof(float)));
    TTX_TRY(ensure(h, h->Qc, (size_t)kp * m * sizeof(float)));
    TTX_TRY(ensure(h, h->Qp, (size_t)kp * m * sizeof(float)));
    {
      KScope ks(h, TTX_K_SMALL, 0);
      launch_copy_cols(a.A, a.m, m, k, P<float>(h->Y), kp, st);
    }
    float* Ya = left_mix(h, a, P<float>(h->Y), kp);
    TTX_TRY(orth(h, Ya, kp, m, k, kp, P<float>(h->Qc), a.stat_dev));
    TTX_TRY(step_tail(h, a, ar, k, kp, 0.0, o));   // Q spans all columns of A: no residual
  } else {
    if (exact) return fail(h, TTX_ERR_UNSUPPORTED, "exact path needs min(m, n) <= TTX_MAX_ELL");
    if (ell > TTX_MAX_ELL) ell = TTX_MAX_ELL;
    if (ell > mn) ell = (int)mn;
    o.path = TTX_PATH_RSVD;
    // DELTA and LEARNED ranks always run the check: their tail estimate uses its residual
    const bool adaptive = (a.flags & TTX_FLAG_ADAPTIVE) || a.pred_ell > 0 || a.rank_mode != TTX_RANK_FIXED;
    // held-out test columns beyond the widest basis, inside the 64-column sketch
    int pt = adaptive ? std::min(8, TTX_MAX_ELL - ell) : 0;
    if (pt < 2) pt = 0;
    const int kp = kp_for(ell + pt);
    SketchPlan sp;
    TTX_TRY(plan_for(h, ar.range, a.A, m, n, kp, &sp));
    TTX_TRY(ensure(h, h->part, (size_t)sp.slabs * m * kp * sizeof(float)));
    TTX_TRY(ensure(h, h->Y, (size_t)m * kp * sizeof(float)));
    TTX_TRY(ensure(h, h->Y2, (size_t)m * kp * sizeof(float)));
    TTX_TRY(ensure(h, h->Qc, (size_t)kp * m * sizeof(float)));
    TTX_TRY(ensure(h, h->Qp, (size_t)kp * m * sizeof(float)));
    TTX_TRY(ensure(h, h->Z, (size_t)n * kp * sizeof(float)));
    if (pt) TTX_TRY(ensure(h, h->Yt, (size_t)m * pt * sizeof(float)));
    struct YfullGuard {   // the full sketch (and the factor taken from it) belongs to this step only
      Handle* h;
      ~YfullGuard() {
        h->Yfull = nullptr;
        h->rprobe_valid = false;
      }
    } yfull_guard{h};
    h->Yfull = nullptr;
    h->rprobe_valid = false;
    if (h->pre_valid && a.A == h->pre_A && a.m == h->pre_m && a.n == h->pre_n && kp == 64 && a.seed == h->pre_seed &&
        a.stream_id == h->pre_stream && ar.range == AR_BF16X3)
      h->Yfull = P<float>(h->Ypre);   // taken in the first step's pass of the sweep
    if (a.rank_mode == TTX_RANK_LEARNED && pt) {
      int pr = 0, pw = 0;
      if (kp == PROBE_K) {
        // the model reads the pivots of this step's own sketch with every Omega column (the range
        // finder's first pass, which the chosen width then reuses), so no separate probe pass
        if (!h->Yfull) {
          float* sc = fp16_scales(h, a, ar, true);
          {
            KScope ks(h, TTX_K_SKETCH, a.m * n * 4);
            launch_sketch(ar.range, sp, a.A, a.m, m, n, nullptr, 0, a.seed, a.stream_id, kp, P<float>(h->part), st,
                          sc);
          }
          {
            KScope ks(h, TTX_K_SMALL, 0);
            launch_reduce_f32(P<float>(h->part), sp.slabs, (int64_t)m * kp, P<float>(h->Y), st);
          }
          TTX_TRY(check_launch(h, "full sketch"));
          o.bytes += a.m * n * 4;
          o.passes += 1;
          h->Yfull = P<float>(h->Y);
        }
        TTX_TRY(predict(h, a, ar, ell, &pr, &pw, o, h->Yfull));
      } else {
        TTX_TRY(predict(h, a, ar, ell, &pr, &pw, o));
      }
      a.pred_ell = pw;
      o.pred_rank = pr;
    }
    // basis widths tried in order: the fixed width; in DELTA mode 16, 32, ... up to r + s; a learned
    // prediction first, then the wider ones of that ladder
    int widths[8];
    int nw = 0;
    if (pt && (a.rank_mode != TTX_RANK_FIXED || a.pred_ell > 0)) {
      int w0 = a.pred_ell > 0 ? std::min(std::max(a.pred_ell, 1), ell) : std::min(16, ell);
      widths[nw++] = w0;
      for (int w = 16; w < ell && nw < 7; w *= 2)
        if (w > w0) widths[nw++] = w;
      if (widths[nw - 1] < ell) widths[nw++] = ell;
      if (a.pred_ell > 0) o.fallback = 0;
    } else {
      widths[nw++] = ell;
    }
    double rel2 = -1.0;
    int wi = 0;
    k = widths[0];
    TTX_TRY(range_sketch(h, a, ar, k, pt, kp, sp, o));
    if (pt) {
      TTX_TRY(range_check(h, a, k, pt, &rel2, o));
      while (rel2 > a.range_tol2 && wi + 1 < nw) {
        if (a.pred_ell > 0) o.fallback = 1;
        k = widths[++wi];
        TTX_TRY(range_sketch(h, a, ar, k, pt, kp, sp, o));
        TTX_TRY(range_check(h, a, k, pt, &rel2, o));
      }
    }
    // power iterations (basicRandSVD_LU :628-683): Z = A^T Q, Y = A Z, Q = orth(Y); adaptive: only
    // while the check fails, q at most (they overwrite Y)
    h->Yfull = nullptr;
    h->rprobe_valid = false;
    int it = 0;
    while (it < a.q && (!pt || rel2 > a.range_tol2)) {
      TTX_TRY(power_iter(h, a, ar, k, kp, sp, o));
      ++it;
      if (pt) TTX_TRY(range_check(h, a, k, pt, &rel2, o));
    }
    if (pt && a.pred_ell > 0 && it > 0) o.fallback = 1;
    o.power_iters = it;
    o.ell = k;
    o.resid = rel2 >= 0 ? std::sqrt(rel2) : -1.0;
    TTX_TRY(step_tail(h, a, ar, k, kp, rel2, o));
  }
  o.kdim = k;
  if (a.sync) {
    TTX_HIP(h, hipMemcpyAsync(h->h_sigma, a.sig_dev, k * sizeof(float), hipMemcpyDeviceToHost, st));
    TTX_HIP(h, hipMemcpyAsync(h->h_status, a.stat_dev, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
    TTX_HIP(h, hipStreamSynchronize(st));
    finish_step_out(h->h_sigma, h->h_status, o);
  }
  return TTX_OK;
}

}  // namespace ttx

// ------------------------------------------------------------------------------------------------
// sweeps
// ------------------------------------------------------------------------------------------------
namespace ttx {

static ttx_status validate(Handle* h, const ttx_params* p, bool tr) {
  if (!p) return fail(h, TTX_ERR_BAD_ARG, "null params");
  if (p->order < 2 || p->order > TTX_MAX_ORDER) return fail(h, TTX_ERR_BAD_ARG, "order out of range");
  for (int k = 0; k < p->order; ++k)
    if (p->dims[k] <= 0) return fail(h, TTX_ERR_BAD_ARG, "non-positive mode size");
  for (int k = 0; k <= p->order; ++k)
    if (p->ranks[k] <= 0) return fail(h, TTX_ERR_BAD_ARG, "non-positive rank");
  if (!tr && (p->ranks[0] != 1 || p->ranks[p->order] != 1)) return fail(h, TTX_ERR_BAD_ARG, "TT ranks need r_0 = r_d = 1");
  if (tr && p->ranks[0] != p->ranks[p->order]) return fail(h, TTX_ERR_BAD_ARG, "TR ranks need r_0 = r_d");
  if (p->oversampling < 0 || p->power_iters < 0) return fail(h, TTX_ERR_BAD_ARG, "negative oversampling / power iterations");
  if (p->precision < 0 || p->precision > 2) return fail(h, TTX_ERR_BAD_ARG, "unknown precision");
  if (p->rank_mode < 0 || p->rank_mode > 2) return fail(h, TTX_ERR_BAD_ARG, "unknown rank mode");
  if (p->method < 0 || p->method > 2) return fail(h, TTX_ERR_BAD_ARG, "unknown method");
  return TTX_OK;
}

static ttx_status input_norm2(Handle* h, const float* A, int64_t count, double* out) {
  const int64_t g = sumsq_grid(count);
  TTX_TRY(ensure(h, h->gpart, (size_t)(g + 1) * sizeof(double)));
  double* part = P<double>(h->gpart);
  launch_sumsq(A, count, part + g, part, g, h->stream);
  TTX_TRY(check_launch(h, "sumsq"));
  TTX_HIP(h, hipMemcpyAsync(h->h_lam, part + g, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  *out = h->h_lam[0];
  return TTX_OK;
}

// the exact 32-row first step of a sweep whose first pass took its Gram, started beside the second
// step: its Gram reduced on the sweep's stream, then the eigensolve (and, for host ranks, the copy of
// its eigenvalues) on estream.  The core is assembled when the second step joins.
static ttx_status side_step0(Handle* h, const StepArgs& a, StepOut& o, bool host_ranks) {
  const int m = (int)a.m;
  hipStream_t st = h->stream;
  if (!h->estream) TTX_HIP(h, hipStreamCreateWithFlags(&h->estream, hipStreamNonBlocking));
  for (hipEvent_t& e : h->eev)
    if (!e) TTX_HIP(h, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!h->h_lam0) TTX_HIP(h, hipHostMalloc((void**)&h->h_lam0, TTX_MAX_ELL * sizeof(double), 0));
  TTX_TRY(ensure(h, h->G0, (size_t)m * m * sizeof(double)));
  TTX_TRY(ensure(h, h->lam0, TTX_MAX_ELL * sizeof(double)));
  TTX_TRY(ensure(h, h->V0, (size_t)m * m * sizeof(double)));
  TTX_TRY(ensure(h, h->keys0, 64 * sizeof(unsigned long long)));
  launch_reduce_f64(P<double>(h->gpart), a.gram_pre, (int64_t)m * m, P<double>(h->G0), st);
  TTX_TRY(check_launch(h, "first step gram"));
  TTX_HIP(h, hipEventRecord(h->eev[0], st));
  TTX_HIP(h, hipStreamWaitEvent(h->estream, h->eev[0], 0));
  launch_syevj(P<double>(h->G0), m, P<double>(h->lam0), P<double>(h->V0), a.stat_dev,
               P<unsigned long long>(h->keys0), h->estream, h->jacobi_max_sweeps);
  TTX_TRY(check_launch(h, "first step eigensolve"));
  if (host_ranks) {
    TTX_HIP(h, hipMemcpyAsync(h->h_lam0, h->lam0.p, m * sizeof(double), hipMemcpyDeviceToHost, h->estream));
    TTX_HIP(h, hipEventRecord(h->eev[1], h->estream));
  }
  memset(o.sigma, 0, sizeof(o.sigma));
  o.path = TTX_PATH_EXACT_M;
  o.ell = 0;
  o.kdim = m;
  o.bytes += a.m * a.n * 4;
  o.passes += 1;
  o.r = m;            // speculated: a full rank ceiling keeps all m singular values unless the ranks
  o.deferred = true;  // are chosen from the eigenvalues (checked when the second step joins)
  return TTX_OK;
}

static ttx_status sweep(Handle* h, const float* d_A, const ttx_params* p, float* const* d_cores, float* h_sigma,
                        ttx_report* rep, bool tr) {
  TTX_TRY(validate(h, p, tr));
  if (!d_A || !d_cores) return fail(h, TTX_ERR_BAD_ARG, "null tensor or core pointers");
  const int d = p->order;
  hipStream_t st = h->stream;
  // TTX_FLAG_BATCH_PLAN: this tensor as a batch decomposes it
  struct ModeGuard {
    Handle* h;
    bool old;
    ~ModeGuard() { h->batch_mode = old; }
  } mode_guard{h, h->batch_mode};
  if (p->flags & TTX_FLAG_BATCH_PLAN) h->batch_mode = true;
  ttx_report R;
  memset(&R, 0, sizeof(R));
  for (int k = 0; k < TTX_MAX_ORDER; ++k) R.norm2_in[k] = -1.0;
  if (h_sigma) memset(h_sigma, 0, sizeof(float) * (d - 1) * TTX_MAX_ELL);
  int64_t total = 1;
  for (int k = 0; k < d; ++k) total *= p->dims[k];
  hipEvent_t ev[TTX_MAX_ORDER + 1];
  const bool timing = (p->flags & TTX_FLAG_TIMING) != 0;
  h->ktime = timing;
  h->evnext = 0;
  h->krecs.clear();
  if (timing)
    for (int k = 0; k < d; ++k) TTX_HIP(h, hipEventCreate(&ev[k]));
  double norm2 = -1.0;
  const double delta_rel = p->eps * p->eps / (double)(d - 1);
  double delta2 = 0.0;
  if (p->rank_mode != TTX_RANK_FIXED) {
    // an exact first step measures ||A||_F^2 as the trace of its Gram: no separate pass
    const int64_t m0 = (int64_t)p->ranks[0] * p->dims[0], n0 = total / p->dims[0];
    const int r0 = tr ? p->ranks[0] * p->ranks[1] : p->ranks[1];
    const bool exact0 = (p->method != TTX_METHOD_RSVD) && m0 <= TTX_MAX_ELL &&
                        ((int64_t)r0 + p->oversampling >= std::min(m0, n0) || p->method == TTX_METHOD_EXACT);
    if (exact0) {
      delta2 = -1.0;
    } else {
      TTX_TRY(input_norm2(h, d_A, total, &norm2));
      delta2 = delta_rel * norm2;
    }
  }

  // per-step sigma / status slots on the device, read back once at the end (FIXED ranks need no
  // host round trip inside the sweep)
  TTX_TRY(ensure(h, h->scratch[3], (size_t)TTX_MAX_ORDER * (TTX_MAX_ELL * sizeof(float) + 4 * sizeof(int))));
  float* sig_all = P<float>(h->scratch[3]);
  int* stat_all = reinterpret_cast<int*>(sig_all + TTX_MAX_ORDER * TTX_MAX_ELL);
  launch_zero(h->scratch[3].p, (int64_t)h->scratch[3].bytes, st);
  const bool host_ranks = p->rank_mode != TTX_RANK_FIXED;
  StepOut outs[TTX_MAX_ORDER];

  const float* cur = d_A;
  int cur_buf = -1;   // which of next[0..1] holds cur (-1: the input)
  int r_prev = 1;
  int64_t n = total;
  const float* pend = nullptr;   // square core of a deferred full-rank step
  int prev_kept = 0;             // singular values kept by the previous step
  bool tr_compact = false;       // the second TR step runs on the rows of its unfolding that can be nonzero
  int tr_r1 = tr ? p->ranks[1] : 0;   // TR: the first step's r1 (DELTA / LEARNED: ceil(R / r0) <= ranks[1])
  int pend_ms = 0;
  R.ranks[0] = p->ranks[0];
  R.ranks[d] = p->ranks[d];
  if (timing) TTX_HIP(h, hipEventRecord(ev[0], st));
  // One read of the tensor for the first two steps: when step 0 is an exact 32-row step whose rank
  // ceiling is full (so it defers its projection) and step 1 is a randomized step with a 64-column
  // sketch, the pass that takes step 0's Gram also takes step 1's sketch (of the same memory viewed
  // as 32 N_1 rows; the deferred core is applied to the sketch afterwards, as for any fused step).
  h->pre_valid = false;
  struct PreGuard {   // the precomputed sketch never outlives this sweep, whatever path leaves it
    Handle* h;
    ~PreGuard() { h->pre_valid = false; }
  } pre_guard{h};
  int64_t gram_pre = 0;
  if (!tr && !h->capture && !h->batch_mode && d >= 3 && p->precision == TTX_PREC_BF16 && !(p->flags & TTX_FLAG_FAST_RANGE) &&
      p->ranks[0] == 1 && p->dims[0] == 32) {
    const int64_t m0 = 32, n0 = total / 32;
    const int64_t r0 = std::min<int64_t>(p->ranks[1], std::min(m0, n0));
    const bool exact0 = (p->method != TTX_METHOD_RSVD) &&
                        (r0 + p->oversampling >= std::min(m0, n0) || p->method == TTX_METHOD_EXACT);
    const int64_t m1 = m0 * p->dims[1], n1 = n0 / p->dims[1];
    const int64_t mn1 = std::min(m1, n1);
    const int64_t r1 = std::min<int64_t>(p->ranks[2], mn1);
    int64_t ell1 = r1 + p->oversampling;
    const bool exact1 = (p->method == TTX_METHOD_EXACT) || (p->method == TTX_METHOD_AUTO && ell1 >= mn1);
    ell1 = std::min<int64_t>(std::min<int64_t>(ell1, TTX_MAX_ELL), mn1);
    const bool adaptive1 = (p->flags & TTX_FLAG_ADAPTIVE) || p->rank_mode != TTX_RANK_FIXED;
    int pt1 = adaptive1 ? (int)std::min<int64_t>(8, TTX_MAX_ELL - ell1) : 0;
    if (pt1 < 2) pt1 = 0;
    if (exact0 && r0 == m0 && !exact1 && kp_for((int)ell1 + pt1) == 64 && m1 % 64 == 0 && m1 <= (1 << 30)) {
      const SketchPlan sp = plan_sketch((int)m1, n1, 64, 0);
      if (sketch_gram0_ok(sp, (int)m1, m1, 64)) {
        const int64_t parts = sketch_gram0_partials(sp);
        TTX_TRY(ensure(h, h->part, (size_t)sp.slabs * m1 * 64 * sizeof(float)));
        TTX_TRY(ensure(h, h->gpart, (size_t)parts * 32 * 32 * sizeof(double)));
        TTX_TRY(ensure(h, h->Ypre, (size_t)m1 * 64 * sizeof(float)));
        // Omega once for the whole pass (a full-machine kernel): the pass itself then only streams A,
        // where generating it per workgroup put the Philox / Box-Muller work (twice, once per row
        // half) on the critical path of every chunk
        TTX_TRY(ensure(h, h->omg, (size_t)omega_chunks_bytes(n1)));
        {
          KScope ks(h, TTX_K_SMALL, 0);
          launch_omega_chunks(p->omega_seed, h->stream_base + 1u, n1, 64, h->omg.p, st);
        }
        {
          KScope ks(h, TTX_K_SKETCH, m1 * n1 * 4);
          launch_sketch_gram0(sp, d_A, m1, (int)m1, n1, p->omega_seed, h->stream_base + 1u, 64, P<float>(h->part),
                              P<double>(h->gpart),
                              st, h->omg.p, (p->flags & TTX_FLAG_TAIL_X3) != 0);
        }
        {
          KScope ks(h, TTX_K_SMALL, 0);
          launch_reduce_f32(P<float>(h->part), sp.slabs, m1 * 64, P<float>(h->Ypre), st);
        }
        TTX_TRY(check_launch(h, "first pass"));
        gram_pre = parts;
        h->pre_valid = true;
        h->pre_A = d_A;
        h->pre_m = m1;
        h->pre_n = n1;
        h->pre_seed = p->omega_seed;
        h->pre_stream = h->stream_base + 1u;
      }
    }
  }
  // the first step's eigensolve beside the second step: when the fused first pass ran and nothing times,
  // captures or runs from the allocation-free arena
  const bool overlap0 = gram_pre > 0 && !timing && !h->capture && !h->arena;
  bool side0 = false, joined0 = false;
  StepArgs a0;
  struct SideGuard {   // no side work outlives the sweep, whatever path leaves it
    Handle* h;
    ~SideGuard() {
      if (h->estream) (void)hipStreamSynchronize(h->estream);
    }
  } side_guard{h};
  // the second step's join: the first step's rank (host ranks: from its eigenvalues), its core on
  // estream, the sweep stream waiting for it; TTX_RETRY when fewer than m singular values are kept
  auto join0 = [&](double* nrm, double* dlt) -> ttx_status {
    joined0 = true;
    StepOut& o0 = outs[0];
    const int m0 = (int)a0.m;
    int r0 = a0.r;
    if (host_ranks) {
      TTX_HIP(h, hipEventSynchronize(h->eev[1]));
      double tot = 0.0;
      for (int i = 0; i < m0; ++i) tot += std::max(h->h_lam0[i], 0.0);
      StepArgs aa = a0;
      if (aa.norm2 < 0) aa.norm2 = tot;
      if (aa.delta2 < 0) aa.delta2 = a0.delta_rel * tot;
      double kept = 0.0;
      r0 = choose_rank(aa, m0, h->h_lam0, tot, 0.0, &kept);
      o0.total2 = tot;
      o0.kept2 = kept;
      // the sweep's bookkeeping of a finished first step
      if (delta2 < 0.0) {
        norm2 = tot;
        delta2 = delta_rel * norm2;
      }
      R.norm2_in[0] = tot;
      norm2 = kept;
    }
    o0.r = r0;
    if (r0 != m0) return TTX_RETRY;
    launch_assemble(nullptr, 0, m0, m0, P<double>(h->lam0), P<double>(h->V0), r0, d_cores[0], nullptr, nullptr, 0,
                    a0.sig_dev, P<unsigned long long>(h->keys0), h->estream);
    TTX_TRY(check_launch(h, "first step core"));
    TTX_HIP(h, hipEventRecord(h->eev[2], h->estream));
    TTX_HIP(h, hipStreamWaitEvent(st, h->eev[2], 0));
    if (nrm) *nrm = norm2;
    if (dlt) *dlt = delta2;
    return TTX_OK;
  };
  for (int k = 0; k < d - 1; ++k) {
    StepArgs a;
    if (k == 0) a.gram_pre = gram_pre;
    const int64_t m = (int64_t)r_prev * p->dims[k];
    n = n / p->dims[k];
    a.A = cur;
    a.m = m;
    a.n = n;
    a.r = (tr && k == 0) ? p->ranks[0] * p->ranks[1] : p->ranks[k + 1];
    if (tr && k == 0) a.rank_multiple = p->ranks[0];
    // the second TR step on the compacted unfolding decides as on its zero-padded rows
    const int64_t m_pad = (tr_compact && k == 1) ? (int64_t)tr_r1 * p->dims[1] : m;
    if (tr_compact && k == 1) a.m_path = m_pad;
    // a rank above min(m, n) cannot be reached: the step keeps min(m, n) (the TR first step keeps its
    // r0 r1 columns, zero-padded, because the relayout needs them)
    if (!(tr && k == 0)) a.r = (int)std::min<int64_t>(a.r, std::min<int64_t>(m_pad, n));
    a.s = p->oversampling;
    a.q = p->power_iters;
    a.seed = p->omega_seed;
    a.stream_id = h->stream_base + (uint32_t)k;
    a.precision = p->precision;
    a.method = p->method;
    a.flags = p->flags;
    a.rank_mode = p->rank_mode;
    a.delta2 = delta2;
    a.delta_rel = delta_rel;
    a.norm2 = norm2;
    a.core = d_cores[k];
    // a deferred full-rank previous step: this step runs on cur itself and its factor reaches only
    // the core (the step beside which the first step's eigensolve runs joins it first)
    a.pend = nullptr;
    a.pend_core = pend;
    a.pend_ms = pend_ms;
    if (side0 && k == 1) a.join = join0;
    a.sig_dev = sig_all + (size_t)k * TTX_MAX_ELL;
    a.sig_prev = (k > 0) ? sig_all + (size_t)(k - 1) * TTX_MAX_ELL : nullptr;
    a.r_prev = prev_kept;
    a.stat_dev = stat_all + 4 * k;
    // the step's own readback is not needed: a DELTA / LEARNED step reads its eigenvalues when it
    // chooses the rank, and the singular values of every step come back once after the sweep
    a.sync = false;
    // range-check tolerance: a quarter of the step's share eps^2 / (d-1) of the squared error budget
    // (Oseledets 2011: the squared step errors add up), relative to ||A_k||_F^2
    // after a deferred full-rank first step ||A_1||^2 = ||A||^2, so delta^2 / ||A_1||^2 is eps^2 / (d-1)
    // itself (taken as such, so a first step still running beside this one gives the same values)
    const bool first_full = (k == 1 && pend);
    if (p->rank_mode == TTX_RANK_FIXED || norm2 <= 0.0 || first_full)
      a.range_tol2 = 0.25 * p->eps * p->eps / (double)(d - 1);
    else
      a.range_tol2 = 0.25 * delta2 / norm2;
    a.rel_budget = (norm2 > 0.0 && !first_full) ? delta2 / norm2 : p->eps * p->eps / (double)(d - 1);
    const bool last = (k == d - 2);
    // A full-rank exact step (r = m, square orthogonal core) that is not the last one is not
    // projected: the next step reads this step's input with the core folded into its small factors.
    const int64_t mn = std::min<int64_t>(m_pad, n);
    const bool exact_m = (a.method != TTX_METHOD_RSVD) && ((int64_t)a.r + a.s >= mn || a.method == TTX_METHOD_EXACT) &&
                         m_pad <= TTX_MAX_ELL;
    const bool can_defer = !last && !(tr && k == 0) && !pend && exact_m && !(tr_compact && k == 1);
    a.defer = can_defer && !host_ranks && a.r == m && !h->batch_mode;
    a.allow_defer = can_defer && host_ranks;
    // TR first step on the exact path (order >= 3): the projection writes the 231 layout itself (no
    // permute pass), and when r0 r1 exceeds the step's rank bound min(r0 r1, N_1) the rows b >= r1c =
    // ceil(bound / r0) are all zero and are not written; the second step then runs on the r1c N_2 rows
    // that can be nonzero (same results: zero rows add nothing to any sketch, Gram or projection) unless
    // its zero-padded unfolding would take the exact short-side path
    if (tr && k == 0 && exact_m && d >= 3) {
      const int64_t r0 = p->ranks[0], r1 = p->ranks[1];
      const int64_t bound = std::min<int64_t>(r0 * r1, m);
      int64_t r1c = (bound + r0 - 1) / r0;
      const int64_t m1 = r1 * p->dims[1], n1 = n * r0 / p->dims[1];
      const int64_t r2 = std::min<int64_t>(p->ranks[2], std::min(m1, n1));
      const bool exact1 = (p->method == TTX_METHOD_EXACT) ||
                          (p->method == TTX_METHOD_AUTO && r2 + p->oversampling >= std::min(m1, n1));
      if (r1c >= r1 || (exact1 && m1 <= TTX_MAX_ELL)) r1c = r1;
      a.tr_r0 = (int)r0;
      a.tr_r1c = (int)r1c;
    }
    float* next = nullptr;
    int wbuf = -1;
    if (last) {
      next = d_cores[d - 1];
    } else if (!a.defer) {
      wbuf = (cur_buf == 0) ? 1 : 0;
      const int64_t rows = a.tr_r1c > 0 ? (int64_t)a.tr_r0 * a.tr_r1c : a.r;
      TTX_TRY(ensure(h, h->next[wbuf], (size_t)rows * n * sizeof(float)));
      next = P<float>(h->next[wbuf]);
    }
    a.next = next;
    // the compacted second TR step writes its (r1c N_2) x r_2 core to scratch; the zero rows are
    // restored into the caller's core after the step
    float* core_compact = nullptr;
    if (tr_compact && k == 1) {
      TTX_TRY(ensure(h, h->scratch[2], (size_t)m * a.r * sizeof(float)));
      core_compact = P<float>(h->scratch[2]);
      a.core = core_compact;
    }
    StepOut& o = outs[k];
    if (k == 0 && overlap0 && (a.defer || a.allow_defer)) {
      TTX_TRY(side_step0(h, a, o, host_ranks));
      side0 = true;
      a0 = a;
    } else {
      const ttx_status s1 = run_step(h, a, o);
      if (s1 == TTX_RETRY && side0 && k == 1) {
        // the first step keeps r0 < m singular values: project it now and run the second step on
        // its remainder (the sweep stream's speculative work is simply superseded)
        side0 = false;
        TTX_HIP(h, hipStreamSynchronize(h->estream));
        StepOut& o0 = outs[0];
        const int m0 = (int)a0.m, r0 = o0.r;
        const int64_t n0 = a0.n;
        const int kpn = kp_for(r0);
        TTX_TRY(ensure(h, h->Qn, (size_t)kpn * m0 * sizeof(float)));
        TTX_TRY(ensure(h, h->next[0], (size_t)r0 * n0 * sizeof(float)));
        const Arith3 ar0 = arith_for(p->precision, p->flags);
        launch_assemble(nullptr, 0, m0, m0, P<double>(h->lam0), P<double>(h->V0), r0, d_cores[0], nullptr,
                        P<float>(h->Qn), kpn, a0.sig_dev, P<unsigned long long>(h->keys0), st);
        launch_proj_short(ar0.accurate, kpn, d_A, m0, m0, n0, P<float>(h->Qn), m0, 1, P<float>(h->next[0]), 0, r0, 0,
                          st);
        TTX_TRY(check_launch(h, "first step projection"));
        o0.deferred = false;
        o0.bytes += (int64_t)m0 * n0 * 4 + (int64_t)r0 * n0 * 4;
        o0.passes += 1;
        R.bytes_alg += (int64_t)m0 * n0 * 4 + (int64_t)r0 * n0 * 4;
        R.passes[0] = o0.passes;
        cur = P<float>(h->next[0]);
        cur_buf = 0;
        r_prev = r0;
        prev_kept = std::min(r0, TTX_MAX_ELL);
        R.ranks[1] = r0;
        pend = nullptr;
        pend_ms = 0;
        h->pre_valid = false;
        n *= p->dims[k];   // the step's iteration divides it again
        outs[1] = StepOut();
        --k;
        continue;
      }
      TTX_TRY(s1);
      if (side0 && k == 1 && !joined0) return fail(h, TTX_ERR_HIP, "internal: the first step was never joined");
    }
    if (core_compact) {
      {
        KScope ks(h, TTX_K_SMALL, 0);
        launch_expand_rows(core_compact, r_prev, tr_r1, p->dims[1], o.r, d_cores[1], st);
      }
      TTX_TRY(check_launch(h, "TR core rows"));
    }
    R.ell[k] = o.ell;
    R.path[k] = o.path;
    R.power_iters[k] = o.power_iters;
    R.bytes_alg += o.bytes;
    R.range_resid[k] = o.resid;
    R.range_checks[k] = o.checks;
    R.passes[k] = o.passes;
    R.fallback[k] = o.fallback;
    R.predicted_rank[k] = o.pred_rank;
    prev_kept = std::min(o.r, TTX_MAX_ELL);
    if (host_ranks && delta2 < 0.0 && o.total2 >= 0.0) {
      norm2 = o.total2;
      delta2 = delta_rel * norm2;
    }
    if (host_ranks) {
      R.norm2_in[k] = (o.total2 >= 0) ? o.total2 : norm2;
      norm2 = o.kept2;
    }
    if (a.defer || o.deferred) {
      pend = d_cores[k];
      pend_ms = (int)m;
      r_prev = o.r;
      R.ranks[k + 1] = o.r;
      if (timing) TTX_HIP(h, hipEventRecord(ev[k + 1], st));
      continue;   // cur stays: the next step reads it with the pending factor
    }
    pend = nullptr;
    pend_ms = 0;
    if (tr && k == 0 && (o.r % p->ranks[0] != 0 || o.r > p->ranks[0] * p->ranks[1]))
      return fail(h, TTX_ERR_BAD_ARG, "TR first step rank mismatch");
    if (tr && k == 0) tr_r1 = o.r / p->ranks[0];
    if (tr && k == 0 && a.tr_r1c > 0) {
      // the projection wrote (r0, r1c, N') as (r1c, N', r0) already; the ring rank joins the columns
      if (o.path != TTX_PATH_EXACT_M) return fail(h, TTX_ERR_HIP, "TR first step: relayout expected on the exact path");
      cur = next;
      cur_buf = wbuf;
      r_prev = o.tr_r1c;
      tr_compact = o.tr_r1c < tr_r1;
      n = n * p->ranks[0];
      R.ranks[1] = tr_r1;
    } else if (tr && k == 0) {
      // (r0, r1, N') -> (r1, N', r0) (trdec6_singleGPU.cu:865-877); the ring rank joins the columns
      const int64_t r0 = p->ranks[0], r1 = tr_r1;
      const int pbuf = (wbuf == 0) ? 1 : 0;
      TTX_TRY(ensure(h, h->next[pbuf], (size_t)o.r * n * sizeof(float)));
      {
        KScope ks(h, TTX_K_PERMUTE, 2 * o.r * n * 4);
        launch_permute_231(next, P<float>(h->next[pbuf]), r0, r1, n, st);
      }
      TTX_TRY(check_launch(h, "permute_231"));
      R.bytes_alg += 2 * o.r * n * 4;
      cur = P<float>(h->next[pbuf]);
      cur_buf = pbuf;
      r_prev = (int)r1;
      n = n * r0;
      R.ranks[1] = (int)r1;
    } else {
      cur = next;
      cur_buf = wbuf;
      r_prev = o.r;
      R.ranks[k + 1] = o.r;
    }
    if (timing) TTX_HIP(h, hipEventRecord(ev[k + 1], st));
  }
  h->pre_valid = false;
  if (h->capture) {
    if (rep) *rep = R;
    return check_launch(h, "batch sweep");
  }
  {
    // sigma and status words are contiguous on the device: one copy into pinned memory
    const float* hs = h->h_sweep;
    const int* hst = reinterpret_cast<const int*>(h->h_sweep + (size_t)TTX_MAX_ORDER * TTX_MAX_ELL);
    TTX_HIP(h, hipMemcpyAsync(h->h_sweep, sig_all, (size_t)TTX_MAX_ORDER * (TTX_MAX_ELL + 4) * sizeof(float),
                              hipMemcpyDeviceToHost, st));
    TTX_HIP(h, hipStreamSynchronize(st));
    double n2 = -1.0;
    for (int k = 0; k < d - 1; ++k) {
      finish_step_out(hs + (size_t)k * TTX_MAX_ELL, hst + 4 * k, outs[k], !host_ranks);
      if (!host_ranks) {
        R.norm2_in[k] = (outs[k].total2 >= 0) ? outs[k].total2 : n2;
        n2 = outs[k].kept2;
      }
    }
  }
  for (int k = 0; k < d - 1; ++k) {
    R.zero_pivots[k] = outs[k].zero_pivots;
    R.jacobi_sweeps[k] = outs[k].sweeps;
    if (h_sigma) memcpy(h_sigma + (size_t)k * TTX_MAX_ELL, outs[k].sigma, sizeof(float) * TTX_MAX_ELL);
  }
  TTX_HIP(h, hipStreamSynchronize(st));
  if (timing) {
    float ms = 0.0f;
    for (int k = 0; k < d - 1; ++k) {
      TTX_HIP(h, hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      R.step_ms[k] = ms;
    }
    TTX_HIP(h, hipEventElapsedTime(&ms, ev[0], ev[d - 1]));
    R.total_ms = ms;
    for (int k = 0; k < d; ++k) (void)hipEventDestroy(ev[k]);
    for (const auto& kr : h->krecs) {
      float kms = 0.0f;
      if (hipEventElapsedTime(&kms, kr.a, kr.b) != hipSuccess) continue;
      R.kernel_ms[kr.kind] += kms;
      R.kernel_launches[kr.kind] += 1;
      R.kernel_bytes[kr.kind] += kr.bytes;
      if (kr.bytes > R.kernel_top_bytes[kr.kind]) {
        R.kernel_top_bytes[kr.kind] = kr.bytes;
        R.kernel_top_ms[kr.kind] = kms;
      }
    }
  }
  h->ktime = false;
  h->krecs.clear();
  if (rep) *rep = R;
  for (int k = 0; k < d - 1; ++k)
    if (outs[k].nonconv)
      return fail(h, TTX_ERR_NO_CONVERGE, "step " + std::to_string(k) + ": the Jacobi eigensolver reached its cap of " +
                                              std::to_string(h->jacobi_max_sweeps) + " sweeps before converging");
  return TTX_OK;
}

// ------------------------------------------------------------------------------------------------
// allocation-free TT sweep: an upper bound of the workspace that sweep() lays out for p, summing every
// buffer request of every step as if each were a fresh piece (the arena never reuses an outgrown
// piece within a call), plus 256 bytes of alignment per piece
// ------------------------------------------------------------------------------------------------
static size_t noalloc_bound(const ttx_params* p) {
  const int d = p->order;
  const size_t F = sizeof(float), D = sizeof(double), E2 = (size_t)TTX_MAX_ELL * TTX_MAX_ELL;
  size_t total_el = 1;
  for (int k = 0; k < d; ++k) total_el *= (size_t)p->dims[k];
  size_t b = 0;
  int pieces = 0;
  auto add = [&](size_t bytes) { b += bytes; ++pieces; };
  add((size_t)TTX_MAX_ORDER * (TTX_MAX_ELL * F + 4 * sizeof(int)));           // sigma / status words
  add((size_t)(sumsq_grid((int64_t)total_el) + 1) * D);                       // input norm
  // the fused first pass of a BF16 sweep with a 32-row first mode
  if (d >= 3 && p->dims[0] == 32) {
    const int64_t m1 = 32 * p->dims[1], n1 = (int64_t)(total_el / (size_t)m1);
    if (m1 <= (1 << 30)) {
      const SketchPlan sp = plan_sketch((int)m1, n1, 64, 0);
      add((size_t)sp.slabs * m1 * 64 * F);
      add((size_t)sketch_gram0_partials(sp) * 32 * 32 * D);
      add((size_t)m1 * 64 * F);
      add((size_t)omega_chunks_bytes(n1));
    }
  }
  int64_t n = (int64_t)total_el;
  for (int k = 0; k < d - 1; ++k) {
    const int64_t m = (int64_t)p->ranks[k] * p->dims[k];
    n /= p->dims[k];
    if (m > ((int64_t)1 << 30)) return (size_t)-1;
    const int64_t r = std::min<int64_t>(p->ranks[k + 1], std::min(m, n));
    const int64_t nb = panel_blocks(m);
    // fixed small buffers: status, sigma, lam, keys, V, G, G2, Rinv, Rprobe, Ur, FP16 scales
    add(16); add(TTX_MAX_ELL * F); add(TTX_MAX_ELL * D); add(64 * 8);
    add(E2 * D); add(E2 * D); add(E2 * D); add(E2 * D); add(E2 * D); add(E2 * F); add(129 * F);
    // sketch partials (the step's own pass and the rank model's probe pass)
    add((size_t)plan_sketch((int)m, n, 64, 0).slabs * m * 64 * F);
    const int64_t np = std::min<int64_t>(n, std::max<int64_t>(std::min<int64_t>(n, 512), n / 16));
    add((size_t)plan_sketch((int)m, np, 64, 0).slabs * m * 64 * F);
    // Gram partials: exact Gram, projection Gram, CholeskyQR panels
    if (m <= TTX_MAX_ELL) add((size_t)gram_small_grid(n) * m * m * D);
    add((size_t)proj_grid(n) * E2 * D);
    add((size_t)nb * E2 * D);
    // tall factors: Y, Y2, Qc, Qp, Qn, Yt; W1, W2; test partials and sums
    for (int i = 0; i < 6; ++i) add((size_t)m * 64 * F);
    add((size_t)m * 64 * D); add((size_t)m * 64 * D);
    add((size_t)nb * 64 * 8 * D); add((size_t)(2 * nb + 2) * D);
    // Z = A^T Q and the next unfolding
    add((size_t)n * 64 * F);
    add((size_t)r * n * F);
  }
  // rank model work buffers (its parameters are the parent handle's)
  add(3 * (size_t)TTX_MLP_PARAMS * F); add((6 * MLP_H + 2 * TTX_MLP_OUT) * F); add(64 * D);
  add(64 * D + (TTX_MLP_IN + TTX_MLP_OUT) * F);
  return b + (size_t)pieces * 256;
}

// ------------------------------------------------------------------------------------------------
// host-streamed TT / TR sweep (ttdec6_singleGPU.cu:540-652, 656-796; trdec6_singleGPU.cu:668-826):
// the tensor stays in host memory; the first step (exact Gram on the N_1 side) reads it in column
// slices, the H2D copy of slice s + 1 on its own stream overlapping the kernels of slice s; its
// remainder (TR: after the 231 relayout) is the device-resident input of the rest of the sweep,
// which is a TT sweep in both cases (TR: over r_1 N_2, N_3, .., N_d r_0, as the device TR sweep).
// ------------------------------------------------------------------------------------------------
static ttx_status sweep_host(Handle* h, const float* h_A, const ttx_params* p, int64_t slice_cols,
                             float* const* d_cores, float* h_sigma, ttx_report* rep, bool tr) {
  TTX_TRY(validate(h, p, tr));
  if (tr && p->rank_mode != TTX_RANK_FIXED) return fail(h, TTX_ERR_UNSUPPORTED, "TR-SVD supports fixed ranks only");
  if (!h_A || !d_cores) return fail(h, TTX_ERR_BAD_ARG, "null tensor or core pointers");
  if (slice_cols < 0) return fail(h, TTX_ERR_BAD_ARG, "negative slice width");
  if (h->capture) return fail(h, TTX_ERR_UNSUPPORTED, "host-streamed sweep inside a batch");
  const int d = p->order;
  if (d < 3) return fail(h, TTX_ERR_UNSUPPORTED, "host-streamed sweep needs order >= 3");
  const int m = (int)p->dims[0];
  if (p->dims[0] > TTX_MAX_ELL) return fail(h, TTX_ERR_UNSUPPORTED, "host-streamed sweep needs N_1 <= TTX_MAX_ELL");
  int64_t n = 1;
  for (int k = 1; k < d; ++k) n *= p->dims[k];
  const int64_t w = slice_cols > 0 ? std::min(slice_cols, n)
                                   : std::min<int64_t>(n, std::max<int64_t>(1024, ((int64_t)64 << 20) / (4 * (int64_t)m)));
  const int64_t S = (n + w - 1) / w;
  hipStream_t st = h->stream;
  const Arith3 ar = arith_for(p->precision, p->flags);
  h->ktime = false;
  h->pre_valid = false;
  if (!h->cstream) TTX_HIP(h, hipStreamCreateWithFlags(&h->cstream, hipStreamNonBlocking));
  for (hipEvent_t& e : h->hev)
    if (!e) TTX_HIP(h, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // one read over PCIe when a device copy of the tensor fits beside the remainder with 2 GB to
  // spare: the slices land in that copy and the projection reads it from HBM
  // TR: the first step keeps its r_0 r_1 columns (zero-padded beyond N_1), as the device sweep
  const int64_t rmax = tr ? (int64_t)p->ranks[0] * p->ranks[1]
                          : std::min<int64_t>({(int64_t)p->ranks[1], (int64_t)m, n});
  if (tr && rmax > TTX_MAX_ELL) return fail(h, TTX_ERR_UNSUPPORTED, "TR first-step rank r_0 r_1 above TTX_MAX_ELL");
  size_t free_b = 0, total_b = 0;
  TTX_HIP(h, hipMemGetInfo(&free_b, &total_b));
  const size_t need = (size_t)(m + (tr ? 2 : 1) * rmax) * n * sizeof(float) + ((size_t)2 << 30);
  const bool resident = !(p->flags & TTX_FLAG_STREAM_TWICE) &&
                        free_b + h->hA.bytes + h->hB.bytes + h->hB2.bytes >= need;
  if (resident) {
    TTX_TRY(ensure(h, h->hA, (size_t)m * n * sizeof(float)));
  } else {
    for (int b = 0; b < 2; ++b) TTX_TRY(ensure(h, h->hs[b], (size_t)m * w * sizeof(float)));
  }
  float* Ad = resident ? P<float>(h->hA) : nullptr;
  const int64_t grid = gram_small_grid(w);
  TTX_TRY(ensure(h, h->gpart, (size_t)grid * m * m * sizeof(double)));
  TTX_TRY(ensure(h, h->hGs, (size_t)S * m * m * sizeof(double)));
  TTX_TRY(ensure(h, h->status, 4 * sizeof(int)));
  TTX_TRY(ensure(h, h->sig, TTX_MAX_ELL * sizeof(float)));
  TTX_TRY(ensure(h, h->lam, TTX_MAX_ELL * sizeof(double)));
  TTX_TRY(ensure(h, h->keys, 64 * sizeof(unsigned long long)));
  TTX_TRY(ensure(h, h->V, TTX_MAX_ELL * TTX_MAX_ELL * sizeof(double)));
  TTX_TRY(ensure(h, h->G, TTX_MAX_ELL * TTX_MAX_ELL * sizeof(double)));
  // the caller's stream may hold earlier work on these buffers: the copies start after it
  TTX_HIP(h, hipEventRecord(h->hev[2], st));
  TTX_HIP(h, hipEventRecord(h->hev[3], st));
  TTX_HIP(h, hipMemsetAsync(h->status.p, 0, 4 * sizeof(int), st));
  // whatever path leaves this function, no copy from the caller's h_A is still in flight
  struct CopyGuard {
    hipStream_t cs;
    ~CopyGuard() { (void)hipStreamSynchronize(cs); }
  } copy_guard{h->cstream};
  // one pass over the slices: copy s into buffer s & 1 once the kernels of slice s - 2 are done, or
  // (resident) into its place in the device copy, after the work queued before this call
  auto stream_pass = [&](auto&& kernels) -> ttx_status {
    for (int64_t s = 0; s < S; ++s) {
      const int b = (int)(s & 1);
      const int64_t c0 = s * w, ws = std::min(w, n - c0);
      float* dst = resident ? Ad + c0 * m : P<float>(h->hs[b]);
      if (!resident || s == 0) TTX_HIP(h, hipStreamWaitEvent(h->cstream, h->hev[2 + b], 0));
      TTX_HIP(h, hipMemcpyAsync(dst, h_A + c0 * m, (size_t)ws * m * sizeof(float), hipMemcpyHostToDevice,
                                h->cstream));
      TTX_HIP(h, hipEventRecord(h->hev[b], h->cstream));
      TTX_HIP(h, hipStreamWaitEvent(st, h->hev[b], 0));
      kernels(s, dst, c0, ws);
      if (!resident) TTX_HIP(h, hipEventRecord(h->hev[2 + b], st));
      TTX_TRY(check_launch(h, "host-streamed slice"));
    }
    return TTX_OK;
  };
  // pass 1: C = A_1 A_1^T as the sum of the slices' Grams (each reduced in FP64, then summed in
  // slice order)
  double* Gs = P<double>(h->hGs);
  TTX_TRY(stream_pass([&](int64_t s, const float* X, int64_t, int64_t ws) {
    launch_gram_small(ar.accurate, X, m, m, ws, P<double>(h->gpart), grid, nullptr, st);
    launch_reduce_f64(P<double>(h->gpart), grid, (int64_t)m * m, Gs + s * m * m, st);
  }));
  launch_reduce_f64(Gs, S, (int64_t)m * m, P<double>(h->G), st);
  int* stat = P<int>(h->status);
  launch_syevj(P<double>(h->G), m, P<double>(h->lam), P<double>(h->V), stat, P<unsigned long long>(h->keys), st,
               h->jacobi_max_sweeps);
  TTX_TRY(check_launch(h, "host-streamed eig"));
  TTX_HIP(h, hipMemcpyAsync(h->h_lam, h->lam.p, m * sizeof(double), hipMemcpyDeviceToHost, st));
  TTX_HIP(h, hipStreamSynchronize(st));
  double total = 0.0;
  for (int i = 0; i < m; ++i) total += std::max(h->h_lam[i], 0.0);
  // rank: fixed (capped at min(N_1, n)), or the delta rule on the exact spectrum with
  // delta^2 = eps^2 ||A||^2 / (d - 1), r_1 as the ceiling
  int r = (int)rmax;
  const double delta2 = p->eps * p->eps * total / (double)(d - 1);
  if (!tr && p->rank_mode != TTX_RANK_FIXED) {
    double tail = total;
    int rr = 0;
    while (rr < r && tail > delta2) tail -= std::max(h->h_lam[rr++], 0.0);
    r = std::max(rr, 1);
  }
  double kept = 0.0;
  for (int i = 0; i < std::min(r, m); ++i) kept += std::max(h->h_lam[i], 0.0);
  const int kpn = kp_for(r);
  TTX_TRY(ensure(h, h->Qn, (size_t)kpn * m * sizeof(float)));
  launch_assemble(nullptr, 0, m, m, P<double>(h->lam), P<double>(h->V), r, d_cores[0], nullptr, P<float>(h->Qn), kpn,
                  P<float>(h->sig), P<unsigned long long>(h->keys), st);
  TTX_TRY(check_launch(h, "host-streamed core"));
  // pass 2: the remainder B = G_1^T A_1 (r x n column-major), slice by slice
  TTX_TRY(ensure(h, h->hB, (size_t)r * n * sizeof(float)));
  float* B = P<float>(h->hB);
  if (resident) {
    launch_proj_short(ar.accurate, kpn, Ad, m, m, n, P<float>(h->Qn), m, 1, B, 0, r, 0, st);
    TTX_TRY(check_launch(h, "host-streamed projection"));
  } else {
    TTX_TRY(stream_pass([&](int64_t, const float* X, int64_t c0, int64_t ws) {
      launch_proj_short(ar.accurate, kpn, X, m, m, ws, P<float>(h->Qn), m, 1, B + c0 * r, 0, r, 0, st);
    }));
  }
  float sig0[TTX_MAX_ELL];
  int st0[4];
  TTX_HIP(h, hipMemcpyAsync(sig0, h->sig.p, sizeof(sig0), hipMemcpyDeviceToHost, st));
  TTX_HIP(h, hipMemcpyAsync(st0, h->status.p, sizeof(st0), hipMemcpyDeviceToHost, st));
  TTX_HIP(h, hipStreamSynchronize(st));
  if (st0[2])
    return fail(h, TTX_ERR_NO_CONVERGE, "step 0: the Jacobi eigensolver reached its cap of " +
                                            std::to_string(h->jacobi_max_sweeps) + " sweeps before converging");
  // the rest of the sweep on B seen as an order d - 1 tensor (r N_2) x N_3 x ... x N_d: its first
  // core (r N_2) x r_2 is the TT core G_2 in place, its later cores G_3 .. G_d.  TR: B relaid out
  // (r_0, r_1, n) -> (r_1, n, r_0) (trdec6_singleGPU.cu:865-877), then the order d - 1 tensor
  // (r_1 N_2) x N_3 x ... x (N_d r_0) whose last core r_{d-1} x N_d r_0 is the TR layout
  const float* Bsub = B;
  const int r1 = tr ? p->ranks[1] : r;
  if (tr) {
    TTX_TRY(ensure(h, h->hB2, (size_t)r * n * sizeof(float)));
    launch_permute_231(B, P<float>(h->hB2), p->ranks[0], p->ranks[1], n, st);
    TTX_TRY(check_launch(h, "host-streamed permute_231"));
    Bsub = P<float>(h->hB2);
  }
  ttx_params q = *p;
  q.order = d - 1;
  q.dims[0] = (int64_t)r1 * p->dims[1];
  for (int k = 1; k < d - 1; ++k) q.dims[k] = p->dims[k + 1];
  if (tr) q.dims[d - 2] *= p->ranks[0];
  q.ranks[0] = 1;
  for (int k = 1; k < d - 1; ++k) q.ranks[k] = p->ranks[k + 1];
  q.ranks[d - 1] = 1;
  if (p->rank_mode != TTX_RANK_FIXED) {
    // the same per-step budget delta^2 for the remaining d - 2 steps: eps'^2 ||B||^2 / (d - 2) = delta^2
    q.eps = (kept > 0.0 && d > 2) ? std::sqrt(delta2 * (double)(d - 2) / kept) : p->eps;
  }
  ttx_report sub;
  float* sub_sigma = h_sigma ? h_sigma + TTX_MAX_ELL : nullptr;
  struct BaseGuard {
    Handle* h;
    ~BaseGuard() { h->stream_base = 0; }
  } base_guard{h};
  h->stream_base = 1;   // the device sweep's step k + 1 draws Omega stream k + 1
  TTX_TRY(sweep(h, Bsub, &q, d_cores + 1, sub_sigma, &sub, false));
  if (h_sigma) memcpy(h_sigma, sig0, sizeof(sig0));
  if (rep) {
    ttx_report R;
    memset(&R, 0, sizeof(R));
    for (int k = 0; k < TTX_MAX_ORDER; ++k) R.norm2_in[k] = -1.0;
    R.ranks[0] = tr ? p->ranks[0] : 1;
    R.ranks[1] = r1;
    for (int k = 2; k < d; ++k) R.ranks[k] = sub.ranks[k - 1];
    R.ranks[d] = tr ? p->ranks[0] : 1;
    R.path[0] = TTX_PATH_EXACT_M;
    R.passes[0] = resident ? 1 : 2;   // reads of the host tensor
    R.jacobi_sweeps[0] = st0[1];
    R.norm2_in[0] = total;
    for (int k = 1; k < d - 1; ++k) {
      R.ell[k] = sub.ell[k - 1];
      R.path[k] = sub.path[k - 1];
      R.power_iters[k] = sub.power_iters[k - 1];
      R.zero_pivots[k] = sub.zero_pivots[k - 1];
      R.jacobi_sweeps[k] = sub.jacobi_sweeps[k - 1];
      R.norm2_in[k] = sub.norm2_in[k - 1];
      R.range_resid[k] = sub.range_resid[k - 1];
      R.range_checks[k] = sub.range_checks[k - 1];
      R.passes[k] = sub.passes[k - 1];
      R.fallback[k] = sub.fallback[k - 1];
      R.predicted_rank[k] = sub.predicted_rank[k - 1];
    }
    R.range_resid[0] = -1.0;
    R.bytes_alg = (int64_t)2 * m * n * 4 + (int64_t)r * n * 4 * (tr ? 3 : 1) + sub.bytes_alg;
    *rep = R;
  }
  return TTX_OK;
}

// ------------------------------------------------------------------------------------------------
// contractions: L (left modes) x R (right modes)
// ------------------------------------------------------------------------------------------------
struct Halves {
  const float* L = nullptr;   // M x K
  const float* R = nullptr;   // K x N
  int64_t M = 0, N = 0, K = 0;
};

static ttx_status contract_halves(Handle* h, int d, const int64_t* dims, const int32_t* ranks,
                                  const float* const* cores, bool tr, Halves& out) {
  hipStream_t st = h->stream;
  const int hsplit = d / 2;   // left modes 0..hsplit-1, right modes hsplit..d-1
  const int64_t r0 = tr ? ranks[0] : 1;
  // left chain
  int64_t P_ = dims[0] * r0;   // rows of the running left factor (ring index after mode 0)
  const float* Lc = cores[0];  // (N0 r0) x r1
  int li = 0;
  for (int k = 1; k < hsplit; ++k) {
    const int64_t cols = dims[k] * ranks[k + 1];
    TTX_TRY(ensure(h, h->scratch[li], (size_t)P_ * cols * sizeof(float)));
    float* Ln = P<float>(h->scratch[li]);
    launch_gemm(P_, cols, ranks[k], 1.0f, Lc, P_, cores[k], ranks[k], 0.0f, Ln, P_, st);
    Lc = Ln;
    P_ *= dims[k];
    li ^= 1;
  }
  // right chain
  int64_t Q_ = dims[d - 1] * r0;
  const float* Rc = cores[d - 1];   // r_{d-1} x (N_{d-1} r0)
  int ri = 2;
  for (int k = d - 2; k >= hsplit; --k) {
    const int64_t rows = (int64_t)ranks[k] * dims[k];
    TTX_TRY(ensure(h, h->scratch[ri], (size_t)rows * Q_ * sizeof(float)));
    float* Rn = P<float>(h->scratch[ri]);
    launch_gemm(rows, Q_, ranks[k + 1], 1.0f, cores[k], rows, Rc, ranks[k + 1], 0.0f, Rn, rows, st);
    Rc = Rn;
    Q_ *= dims[k];
    ri = (ri == 2) ? 3 : 2;
  }
  TTX_TRY(check_launch(h, "contract chains"));
  const int64_t rh = ranks[hsplit];
  if (!tr) {
    out.L = Lc; out.R = Rc; out.M = P_; out.N = Q_; out.K = rh;
    return TTX_OK;
  }
  // TR: move the ring index next to the bond index on both sides
  const int64_t N0 = dims[0];
  const int64_t Pp = P_ / (N0 * r0);
  const int64_t Qp = Q_ / r0;
  // left: (N0, r0, Pp, rh) -> (N0, Pp, r0, rh); right: (rh, Qp, r0) -> (r0, rh, Qp)
  DevBuf& lb = h->scratch[li];
  TTX_TRY(ensure(h, lb, (size_t)P_ * rh * sizeof(float)));
  {
    const int64_t dl[4] = {N0, r0, Pp, rh};
    const int pl[4] = {0, 2, 1, 3};
    launch_permute(Lc, P<float>(lb), 4, dl, pl, st);
  }
  DevBuf& rb = h->scratch[ri];
  TTX_TRY(ensure(h, rb, (size_t)rh * Q_ * sizeof(float)));
  {
    const int64_t dr[3] = {rh, Qp, r0};
    const int pr[3] = {2, 0, 1};
    launch_permute(Rc, P<float>(rb), 3, dr, pr, st);
  }
  TTX_TRY(check_launch(h, "TR permutes"));
  out.L = P<float>(lb); out.R = P<float>(rb); out.M = N0 * Pp; out.N = Qp; out.K = r0 * rh;
  return TTX_OK;
}

// ------------------------------------------------------------------------------------------------
// TT rounding (Oseledets 2011, TT-rounding, Algorithm 2): a consumer of the sweep's cores that
// recompresses a TT (for instance the sum of two TTs, whose ranks add up) to the smallest ranks whose
// relative error is within eps.  Right to left, core k (r_{k-1} x N_k r_k) becomes Q^T with Q the
// orthonormal basis of its transpose (shifted CholeskyQR2, which stays orthonormal on a rank-deficient
// core), and R^T = M Q moves into core k - 1.  Left to right, each core's unfolding then takes one
// exact step of the sweep with the delta rule (delta^2 = eps^2 ||A||^2 / (d - 1), ||A|| being the
// first core's norm once the others are right-orthonormal): its core is U_r and S V^T moves into
// core k + 1.  Ranks at most TTX_MAX_ELL; FP32 data, FP64 solves.
// ------------------------------------------------------------------------------------------------
static ttx_status tt_round(Handle* h, int d, const int64_t* dims, const int32_t* rin, const float* const* cin,
                           double eps, const int32_t* rmax, float* const* cout, int32_t* rout, float* h_sigma) {
  hipStream_t st = h->stream;
  if (d < 2 || d > TTX_MAX_ORDER || !(eps >= 0.0)) return fail(h, TTX_ERR_BAD_ARG, "round: bad order or eps");
  if (rin[0] != 1 || rin[d] != 1) return fail(h, TTX_ERR_BAD_ARG, "round: TT ranks need r_0 = r_d = 1");
  for (int k = 0; k < d; ++k) {
    if (dims[k] <= 0 || rin[k + 1] <= 0 || !cin[k] || !cout[k]) return fail(h, TTX_ERR_BAD_ARG, "round: bad core");
    if (rin[k + 1] > TTX_MAX_ELL) return fail(h, TTX_ERR_UNSUPPORTED, "round: rank above TTX_MAX_ELL");
    if (k > 0 && rin[k] > dims[k] * rin[k + 1])
      return fail(h, TTX_ERR_UNSUPPORTED, "round: r_{k-1} above N_k r_k (reduce the core first)");
  }
  // working copies W[k] (each sized for its input ranks; ranks only shrink), then a contraction buffer
  std::vector<size_t> off(d + 1, 0);
  size_t tmp_floats = 0;
  for (int k = 0; k < d; ++k) {
    const size_t e = (size_t)rin[k] * dims[k] * rin[k + 1];
    off[k + 1] = off[k] + ((e + 63) & ~(size_t)63);
    tmp_floats = std::max(tmp_floats, e);
  }
  TTX_TRY(ensure(h, h->rwork, (off[d] + tmp_floats + 2 * 64 * 64) * sizeof(float)));
  float* base = P<float>(h->rwork);
  std::vector<float*> W(d);
  for (int k = 0; k < d; ++k) {
    W[k] = base + off[k];
    TTX_HIP(h, hipMemcpyAsync(W[k], cin[k], (size_t)rin[k] * dims[k] * rin[k + 1] * sizeof(float),
                              hipMemcpyDeviceToDevice, st));
  }
  float* T = base + off[d];               // contraction output
  float* RT = T + tmp_floats;             // R^T (r x r) of the orthogonalization
  std::vector<int> r(rin, rin + d + 1);
  TTX_TRY(ensure(h, h->status, 4 * sizeof(int)));
  TTX_HIP(h, hipMemsetAsync(h->status.p, 0, 4 * sizeof(int), st));
  // right to left: core k = Q^T, core k - 1 times R^T
  for (int k = d - 1; k >= 1; --k) {
    const int rk = r[k];
    const int64_t L = dims[k] * (int64_t)r[k + 1];
    const int kp = kp_for(rk);
    TTX_TRY(ensure(h, h->Qc, (size_t)kp * L * sizeof(float)));
    // core k, r_{k-1} x L column-major, is M^T as an L x r_{k-1} row-major matrix
    TTX_TRY(orth(h, W[k], rk, (int)L, rk, kp, P<float>(h->Qc), P<int>(h->status)));
    launch_gemm(rk, rk, L, 1.0f, W[k], rk, P<float>(h->Qc), L, 0.0f, RT, rk, st);   // R^T = M Q
    const int64_t tdims[2] = {L, rk};
    const int tperm[2] = {1, 0};
    launch_permute(P<float>(h->Qc), W[k], 2, tdims, tperm, st);                    // core k = Q^T
    const int64_t mk = (int64_t)r[k - 1] * dims[k - 1];
    launch_gemm(mk, rk, rk, 1.0f, W[k - 1], mk, RT, rk, 0.0f, T, mk, st);
    TTX_HIP(h, hipMemcpyAsync(W[k - 1], T, (size_t)mk * rk * sizeof(float), hipMemcpyDeviceToDevice, st));
    TTX_TRY(check_launch(h, "round: orthogonalization"));
  }
  double norm2 = 0.0;
  TTX_TRY(input_norm2(h, W[0], (int64_t)dims[0] * r[1], &norm2));
  const double delta2 = eps * eps * norm2 / (double)(d - 1);
  // left to right: one exact step with the delta rule per core
  TTX_TRY(ensure(h, h->scratch[3], (size_t)TTX_MAX_ELL * (TTX_MAX_ELL + 4) * sizeof(float)));
  for (int k = 0; k < d - 1; ++k) {
    StepArgs a;
    a.A = W[k];
    a.m = (int64_t)r[k] * dims[k];
    a.n = r[k + 1];
    const int cap = rmax ? std::max(1, rmax[k + 1]) : r[k + 1];
    a.r = (int)std::min<int64_t>(cap, std::min<int64_t>(a.m, a.n));
    a.s = 0;
    a.q = 0;
    a.precision = TTX_PREC_FP32;
    a.method = TTX_METHOD_EXACT;
    a.rank_mode = TTX_RANK_DELTA;
    a.delta2 = delta2;
    a.precise = true;
    a.core = cout[k];
    a.next = T;
    a.sig_dev = P<float>(h->scratch[3]);
    a.stat_dev = reinterpret_cast<int*>(P<float>(h->scratch[3]) + TTX_MAX_ELL);
    TTX_HIP(h, hipMemsetAsync(a.stat_dev, 0, 4 * sizeof(int), st));
    a.sync = true;
    StepOut o;
    TTX_TRY(run_step(h, a, o));
    if (o.nonconv)
      return fail(h, TTX_ERR_NO_CONVERGE, "round: the Jacobi eigensolver reached its sweep cap at bond " +
                                              std::to_string(k + 1));
    if (h_sigma)
      for (int i = 0; i < TTX_MAX_ELL; ++i) h_sigma[(size_t)k * TTX_MAX_ELL + i] = i < o.kdim ? o.sigma[i] : 0.0f;
    const int rn = o.r;
    // core k + 1 = (S V^T) times core k + 1 (r[k+1] x N_{k+1} r_{k+2})
    const int64_t nk = dims[k + 1] * (int64_t)r[k + 2];
    float* dst = (k + 1 == d - 1) ? cout[d - 1] : W[k + 1];
    if (k + 1 == d - 1) {
      launch_gemm(rn, nk, r[k + 1], 1.0f, T, rn, W[k + 1], r[k + 1], 0.0f, dst, rn, st);
    } else {
      // the product may not overwrite its operand: behind S V^T in the contraction buffer (which runs on
      // into the R^T space, free in this sweep)
      float* prod = T + (size_t)rn * r[k + 1];
      if ((size_t)rn * r[k + 1] + (size_t)rn * nk > tmp_floats + 2 * 64 * 64)
        return fail(h, TTX_ERR_HIP, "internal: round contraction buffer");
      launch_gemm(rn, nk, r[k + 1], 1.0f, T, rn, W[k + 1], r[k + 1], 0.0f, prod, rn, st);
      TTX_HIP(h, hipMemcpyAsync(W[k + 1], prod, (size_t)rn * nk * sizeof(float), hipMemcpyDeviceToDevice, st));
    }
    TTX_TRY(check_launch(h, "round: truncation"));
    r[k + 1] = rn;
  }
  TTX_HIP(h, hipStreamSynchronize(st));
  for (int k = 0; k <= d; ++k) rout[k] = r[k];
  return TTX_OK;
}

static ttx_status rel_error(Handle* h, int d, const int64_t* dims, const int32_t* ranks, const float* const* cores,
                            const float* A, bool tr, double* err) {
  Halves hv;
  TTX_TRY(contract_halves(h, d, dims, ranks, cores, tr, hv));
  const int64_t nb = gemm_resid_blocks(hv.M, hv.N);
  TTX_TRY(ensure(h, h->gpart, (size_t)nb * 2 * sizeof(double)));
  launch_gemm_resid(hv.M, hv.N, hv.K, hv.L, hv.M, hv.R, hv.K, A, hv.M, P<double>(h->gpart), h->stream);
  TTX_TRY(check_launch(h, "residual"));
  std::vector<double> buf((size_t)nb * 2);
  TTX_HIP(h, hipMemcpyAsync(buf.data(), h->gpart.p, nb * 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  double r2 = 0.0, x2 = 0.0;
  for (int64_t i = 0; i < nb; ++i) { r2 += buf[2 * i]; x2 += buf[2 * i + 1]; }
  *err = (x2 > 0) ? std::sqrt(r2 / x2) : std::sqrt(r2);
  return TTX_OK;
}

static ttx_status full(Handle* h, int d, const int64_t* dims, const int32_t* ranks, const float* const* cores,
                       float* out, bool tr) {
  Halves hv;
  TTX_TRY(contract_halves(h, d, dims, ranks, cores, tr, hv));
  launch_gemm(hv.M, hv.N, hv.K, 1.0f, hv.L, hv.M, hv.R, hv.K, 0.0f, out, hv.M, h->stream);
  TTX_TRY(check_launch(h, "full"));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  return TTX_OK;
}

static int64_t core_elems(int d, const int64_t* dims, const int32_t* ranks, bool tr, int k) {
  if (!tr) return (int64_t)ranks[k] * dims[k] * ranks[k + 1];
  if (k == 0) return dims[0] * (int64_t)ranks[0] * ranks[1];
  if (k == d - 1) return (int64_t)ranks[d - 1] * dims[d - 1] * ranks[0];
  return (int64_t)ranks[k] * dims[k] * ranks[k + 1];
}

// glibc's rand() after srand(seed) (the generator of the reference's inputs, generalTTTensor /
// generalTRTensor, ttdec6_singleGPU.cu:185-271, trdec6_singleGPU.cu:231-329): srandom_r's Park-Miller
// sequence r_0 = seed, r_i = 16807 r_{i-1} mod (2^31 - 1) for i < 31, then r_i = r_{i-31} for
// i = 31..33 and r_i = r_{i-3} + r_{i-31} mod 2^32 beyond (random_r, TYPE_3 additive feedback);
// output k is r_{k+344} >> 1 (the first 310 outputs are discarded by srandom_r).
class GlibcRand {
 public:
  explicit GlibcRand(uint32_t seed) {
    int64_t w = seed == 0 ? 1 : (int32_t)seed;
    s_[0] = (uint32_t)w;
    for (int i = 1; i < 31; ++i) {
      // Schrage's form of 16807 w mod (2^31 - 1), as srandom_r computes it on signed 32-bit words
      const int64_t hi = w / 127773, lo = w % 127773;
      w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      s_[i] = (uint32_t)w;
    }
    // r_31..r_33 repeat r_0..r_2 (they sit in the same slots); the additive outputs start at r_34
    i_ = 34;
    for (int k = 0; k < 310; ++k) (void)next();
  }
  // RAND_MAX = 2^31 - 1
  uint32_t next() {
    const int slot = (int)(i_ % 31);
    s_[slot] += s_[(i_ - 3) % 31];   // r_i = r_{i-31} + r_{i-3}: slot i mod 31 holds r_{i-31}
    ++i_;
    return s_[slot] >> 1;
  }

 private:
  uint32_t s_[31];
  int64_t i_;
};

// entry e of a reference core: (float)(rand() / RAND_MAX) * 2.0 - 1.0, rounded to float
// (the reference's expression: the quotient is cast to float before the affine map in double)
static float glibc_uniform(GlibcRand& g) {
  const float u = (float)((double)g.next() / 2147483647.0);
  return (float)((double)u * 2.0 - 1.0);
}

// cores from glibc's rand() in the reference's order: G_{d-1}, G_d, then G_{d-2} down to G_1
// (order 6: G5, G6, G4, G3, G2, G1), each filled in storage order
static void glibc_cores(int d, const std::vector<int64_t>& sz, uint32_t seed, std::vector<std::vector<float>>& out) {
  GlibcRand g(seed);
  out.assign(d, {});
  std::vector<int> order;
  if (d >= 2) {
    order.push_back(d - 2);
    order.push_back(d - 1);
    for (int k = d - 3; k >= 0; --k) order.push_back(k);
  }
  for (int k : order) {
    out[k].resize((size_t)sz[k]);
    for (int64_t e = 0; e < sz[k]; ++e) out[k][(size_t)e] = glibc_uniform(g);
  }
}

// source: 0 = Philox normal cores (std_, noise_rel), 1 = the reference's glibc uniform cores (no noise)
static ttx_status generate(Handle* h, int d, const int64_t* dims, const int32_t* ranks, float std_, float noise_rel,
                           uint64_t seed, float* A, float* const* d_cores, bool tr, int source = 0) {
  if (d < 2 || d > TTX_MAX_ORDER || !A) return fail(h, TTX_ERR_BAD_ARG, "generate: bad order or null output");
  for (int k = 0; k < d; ++k)
    if (dims[k] <= 0) return fail(h, TTX_ERR_BAD_ARG, "generate: bad mode size");
  for (int k = 0; k <= d; ++k)
    if (ranks[k] <= 0) return fail(h, TTX_ERR_BAD_ARG, "generate: bad rank");
  // cores in one scratch block unless the caller wants them
  std::vector<int64_t> sz(d), off(d + 1, 0);
  for (int k = 0; k < d; ++k) {
    sz[k] = core_elems(d, dims, ranks, tr, k);
    off[k + 1] = off[k] + ((sz[k] + 63) / 64) * 64;
  }
  DevBuf cb;
  std::vector<float*> cp(d);
  if (!d_cores) {
    if (hipMalloc(&cb.p, off[d] * sizeof(float)) != hipSuccess) return fail(h, TTX_ERR_OOM, "generate: cores");
    for (int k = 0; k < d; ++k) cp[k] = P<float>(cb) + off[k];
  } else {
    for (int k = 0; k < d; ++k) cp[k] = d_cores[k];
  }
  if (source == 1) {
    std::vector<std::vector<float>> hc;
    glibc_cores(d, sz, (uint32_t)seed, hc);
    for (int k = 0; k < d; ++k) {
      const hipError_t e = hipMemcpyAsync(cp[k], hc[k].data(), (size_t)sz[k] * sizeof(float), hipMemcpyHostToDevice,
                                          h->stream);
      if (e != hipSuccess) {
        if (cb.p) (void)hipFree(cb.p);
        return fail(h, TTX_ERR_HIP, std::string("generate: core copy: ") + hipGetErrorString(e));
      }
    }
    // the pageable sources must outlive the copies
    const hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
      if (cb.p) (void)hipFree(cb.p);
      return fail(h, TTX_ERR_HIP, std::string("generate: core copy: ") + hipGetErrorString(e));
    }
    noise_rel = 0.0f;
  } else {
    for (int k = 0; k < d; ++k)
      launch_fill_normal(cp[k], sz[k], seed, (uint32_t)((tr ? 0x200 : 0x100) + k), std_, h->stream);
  }
  ttx_status s = full(h, d, dims, ranks, cp.data(), A, tr);
  if (s == TTX_OK && noise_rel != 0.0f) {
    int64_t total = 1;
    for (int k = 0; k < d; ++k) total *= dims[k];
    double n2 = 0.0;
    s = input_norm2(h, A, total, &n2);
    if (s == TTX_OK) {
      const float rms = (float)std::sqrt(n2 / (double)total);
      const float coef = noise_rel * rms;
      launch_add_noise(A, total, seed, 0xFFFFu, coef, h->stream);
      s = check_launch(h, "noise");
      if (s == TTX_OK) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) s = fail(h, TTX_ERR_HIP, hipGetErrorString(e));
      }
    }
  }
  if (cb.p) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(cb.p);
  }
  return s;
}

}  // namespace ttx

// ------------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------------
using ttx::Handle;

static Handle* H(ttx_handle h) { return reinterpret_cast<Handle*>(h); }

extern "C" {

int32_t ttx_version(void) { return 100; }

const char* ttx_status_string(int32_t s) {
  switch (s) {
    case TTX_OK: return "ok";
    case TTX_ERR_BAD_ARG: return "bad argument";
    case TTX_ERR_OOM: return "out of device memory";
    case TTX_ERR_NO_CONVERGE: return "no convergence";
    case TTX_RESIDUAL_FALLBACK_USED: return "ok (adaptive fallback used)";
    case TTX_ERR_HIP: return "HIP runtime error";
    case TTX_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
  }
}

static void free_handle(Handle* h, bool own_stream);

ttx_status ttx_create(ttx_handle* out, void* stream) {
  if (!out) return TTX_ERR_BAD_ARG;
  Handle* h = new Handle();
  h->stream = reinterpret_cast<hipStream_t>(stream);
  if (hipGetDevice(&h->dev) != hipSuccess ||
      hipHostMalloc((void**)&h->h_status, 16 * sizeof(int), 0) != hipSuccess ||
      hipHostMalloc((void**)&h->h_sigma, TTX_MAX_ELL * sizeof(float), 0) != hipSuccess ||
      hipHostMalloc((void**)&h->h_lam, TTX_MAX_ELL * sizeof(double), 0) != hipSuccess ||
      hipHostMalloc((void**)&h->h_aux, 16 * sizeof(double), 0) != hipSuccess ||
      hipHostMalloc((void**)&h->h_sweep, (size_t)TTX_MAX_ORDER * (TTX_MAX_ELL + 4) * sizeof(float), 0) != hipSuccess) {
    (void)hipGetLastError();
    delete h;
    *out = nullptr;
    return TTX_ERR_HIP;
  }
  (void)ttx::num_cus();
  // the allocation-free entry point's scratch array: its device address on this device, resolved once
  // per device at set-up; and the shadow handle that runs those calls (host memory only)
  if (ttx::noalloc_init(h->dev) != hipSuccess) {
    (void)hipGetLastError();
    free_handle(h, false);
    *out = nullptr;
    return TTX_ERR_HIP;
  }
  h->noalloc = new Handle();
  // TTX_JACOBI_MAX_SWEEPS (1..30): a lower sweep cap, so callers and tests can see TTX_ERR_NO_CONVERGE
  if (const char* js = getenv("TTX_JACOBI_MAX_SWEEPS")) {
    const int v = atoi(js);
    if (v >= 1 && v <= ttx::kJacobiMaxSweeps) h->jacobi_max_sweeps = v;
  }
  *out = reinterpret_cast<ttx_handle>(h);
  return TTX_OK;
}

// every device workspace buffer of a handle
static std::vector<ttx::DevBuf*> handle_bufs(Handle* h) {
  return {&h->part, &h->Y, &h->Y2, &h->Qc, &h->Qp, &h->Qn, &h->Z, &h->W1, &h->W2, &h->G, &h->G2,
          &h->Rinv, &h->lam,
          &h->V, &h->Ur, &h->gpart, &h->status, &h->next[0], &h->next[1], &h->sig, &h->keys,
          &h->scratch[0], &h->scratch[1], &h->scratch[2], &h->scratch[3], &h->Yt, &h->tpart,
          &h->tout, &h->mP, &h->mM1, &h->mM2, &h->mG, &h->mAct, &h->mPart, &h->mIO,
          &h->f16s, &h->Ypre, &h->omg, &h->Rprobe, &h->hs[0], &h->hs[1], &h->hGs, &h->hB, &h->hB2, &h->hA,
          &h->G0, &h->lam0, &h->V0, &h->keys0, &h->barena, &h->bblob, &h->rwork};
}

static void free_handle(Handle* h, bool own_stream) {
  (void)hipStreamSynchronize(h->stream);
  // the allocation-free shadow owns nothing: its buffers are pieces of the static scratch array or
  // the parent's, its host words are the parent's
  if (h->noalloc) delete h->noalloc;
  h->noalloc = nullptr;
  for (Handle* sub : h->subs) free_handle(sub, true);
  h->subs.clear();
  h->gplan.reset();
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->bstream) (void)hipStreamDestroy(h->bstream);
  if (h->cstream) {
    (void)hipStreamSynchronize(h->cstream);
    (void)hipStreamDestroy(h->cstream);
  }
  for (hipEvent_t e : h->hev)
    if (e) (void)hipEventDestroy(e);
  if (h->estream) {
    (void)hipStreamSynchronize(h->estream);
    (void)hipStreamDestroy(h->estream);
  }
  for (hipEvent_t e : h->eev)
    if (e) (void)hipEventDestroy(e);
  for (auto* b : handle_bufs(h))
    if (b->p) (void)hipFree(b->p);
  for (float* c : h->last_cores) (void)hipFree(c);
  for (hipEvent_t e : h->evpool) (void)hipEventDestroy(e);
  if (h->h_status) (void)hipHostFree(h->h_status);
  if (h->h_sigma) (void)hipHostFree(h->h_sigma);
  if (h->h_lam) (void)hipHostFree(h->h_lam);
  if (h->h_aux) (void)hipHostFree(h->h_aux);
  if (h->h_sweep) (void)hipHostFree(h->h_sweep);
  if (h->h_lam0) (void)hipHostFree(h->h_lam0);
  if (own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

ttx_status ttx_destroy(ttx_handle hh) {
  Handle* h = H(hh);
  if (!h) return TTX_OK;
  free_handle(h, false);
  return TTX_OK;
}

ttx_status ttx_set_stream(ttx_handle h, void* stream) {
  if (!h) return TTX_ERR_BAD_ARG;
  H(h)->stream = reinterpret_cast<hipStream_t>(stream);
  return TTX_OK;
}

const char* ttx_last_error(ttx_handle h) { return h ? H(h)->err.c_str() : "null handle"; }

void ttx_params_init(ttx_params* p, int32_t order, const int64_t* dims, const int32_t* ranks) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->order = order;
  for (int k = 0; k < order && k < TTX_MAX_ORDER; ++k) p->dims[k] = dims ? dims[k] : 1;
  for (int k = 0; k <= order && k <= TTX_MAX_ORDER; ++k) p->ranks[k] = ranks ? ranks[k] : 1;
  p->oversampling = 10;
  p->power_iters = 2;
  p->omega_seed = 1234;
  p->precision = TTX_PREC_BF16;
  p->method = TTX_METHOD_AUTO;
  p->rank_mode = TTX_RANK_FIXED;
  p->eps = 1e-4;
  p->flags = 0;
}

int64_t ttx_core_size(const ttx_params* p, int32_t tr, int32_t k) {
  if (!p || k < 0 || k >= p->order) return 0;
  return ttx::core_elems(p->order, p->dims, p->ranks, tr != 0, k);
}

ttx_status ttx_tt_svd(ttx_handle h, const float* d_A, const ttx_params* p, float* const* d_cores, float* h_sigma,
                      ttx_report* rep) {
  if (!h) return TTX_ERR_BAD_ARG;
  return ttx::sweep(H(h), d_A, p, d_cores, h_sigma, rep, false);
}

ttx_status ttx_tt_svd_noalloc(ttx_handle hh, const float* d_A, const ttx_params* p, float* const* d_cores,
                              float* h_sigma, ttx_report* rep) {
  Handle* h = H(hh);
  if (!h) return TTX_ERR_BAD_ARG;
  Handle* s = h->noalloc;
  if (!s) return ttx::fail(h, TTX_ERR_HIP, "ttx_tt_svd_noalloc: handle without its scratch set-up");
  TTX_TRY(ttx::validate(h, p, false));
  if (p->flags & TTX_FLAG_TIMING)
    return ttx::fail(h, TTX_ERR_UNSUPPORTED, "ttx_tt_svd_noalloc: no per-kernel timing (its events are created per call)");
  const size_t need = ttx::noalloc_bound(p);
  if (need > (size_t)TTX_NOALLOC_SCRATCH_BYTES)
    return ttx::fail(h, TTX_ERR_UNSUPPORTED,
                     "ttx_tt_svd_noalloc: the workspace of this decomposition (up to " + std::to_string(need) +
                         " bytes) exceeds the scratch maximum of " + std::to_string((size_t)TTX_NOALLOC_SCRATCH_BYTES) +
                         " bytes; use ttx_tt_svd");
  unsigned char* base = static_cast<unsigned char*>(ttx::noalloc_scratch(h->dev));
  if (!base) return ttx::fail(h, TTX_ERR_HIP, "ttx_tt_svd_noalloc: scratch address not resolved");
  // the shadow runs on the parent's stream with the parent's host words and rank model; every buffer
  // is a fresh piece of this device's scratch array (the previous call's pieces are forgotten, not freed)
  for (ttx::DevBuf* b : handle_bufs(s)) *b = ttx::DevBuf();
  s->stream = h->stream;
  s->dev = h->dev;
  s->h_status = h->h_status;
  s->h_sigma = h->h_sigma;
  s->h_lam = h->h_lam;
  s->h_aux = h->h_aux;
  s->h_sweep = h->h_sweep;
  s->jacobi_max_sweeps = h->jacobi_max_sweeps;
  s->mP = h->mP;
  s->mlp_ready = h->mlp_ready;
  s->arena = base;
  s->arena_cap = TTX_NOALLOC_SCRATCH_BYTES;
  s->arena_used = 0;
  const ttx_status st = ttx::sweep(s, d_A, p, d_cores, h_sigma, rep, false);
  if (st != TTX_OK) h->err = s->err;
  return st;
}

ttx_status ttx_tr_svd(ttx_handle h, const float* d_A, const ttx_params* p, float* const* d_cores, float* h_sigma,
                      ttx_report* rep) {
  if (!h) return TTX_ERR_BAD_ARG;
  return ttx::sweep(H(h), d_A, p, d_cores, h_sigma, rep, true);
}

ttx_status ttx_tt_svd_host(ttx_handle h, const float* h_A, const ttx_params* p, int64_t slice_cols,
                           float* const* d_cores, float* h_sigma, ttx_report* rep) {
  if (!h) return TTX_ERR_BAD_ARG;
  return ttx::sweep_host(H(h), h_A, p, slice_cols, d_cores, h_sigma, rep, false);
}

ttx_status ttx_tr_svd_host(ttx_handle h, const float* h_A, const ttx_params* p, int64_t slice_cols,
                           float* const* d_cores, float* h_sigma, ttx_report* rep) {
  if (!h) return TTX_ERR_BAD_ARG;
  return ttx::sweep_host(H(h), h_A, p, slice_cols, d_cores, h_sigma, rep, true);
}

ttx_status ttx_step(ttx_handle hh, const float* d_A, int64_t m, int64_t n, int32_t r, int32_t s, int32_t q,
                    uint64_t seed, uint32_t stream_id, int32_t precision, int32_t method, float* d_core, float* d_next,
                    float* h_sigma, int32_t* path_out) {
  Handle* h = H(hh);
  if (!h || !d_A || !d_core || !d_next || m <= 0 || n <= 0 || r <= 0 || s < 0 || q < 0) {
    if (h) h->err = "ttx_step: bad argument";
    return TTX_ERR_BAD_ARG;
  }
  ttx::StepArgs a;
  a.A = d_A; a.m = m; a.n = n; a.r = r; a.s = s; a.q = q; a.seed = seed; a.stream_id = stream_id;
  a.precision = precision; a.method = method; a.core = d_core; a.next = d_next;
  ttx::StepOut o;
  ttx_status st = ttx::run_step(h, a, o);
  if (st != TTX_OK) return st;
  if (h_sigma) memcpy(h_sigma, o.sigma, sizeof(float) * TTX_MAX_ELL);
  if (path_out) *path_out = o.path;
  if (o.nonconv) return ttx::fail(h, TTX_ERR_NO_CONVERGE, "ttx_step: the Jacobi eigensolver did not converge");
  return TTX_OK;
}

ttx_status ttx_rsvd(ttx_handle hh, const float* d_A, int64_t m, int64_t n, int32_t k, int32_t s, int32_t q,
                    uint64_t seed, int32_t precision, float* d_U, float* d_S, float* d_SVT) {
  Handle* h = H(hh);
  if (!h || !d_A || !d_U || !d_SVT || m <= 0 || n <= 0 || k <= 0 || s < 0 || q < 0) {
    if (h) h->err = "ttx_rsvd: bad argument";
    return TTX_ERR_BAD_ARG;
  }
  const int ell = k + s;
  if (ell > TTX_MAX_ELL || ell > std::min(m, n)) {
    h->err = "ttx_rsvd: k + s must not exceed min(m, n) or TTX_MAX_ELL";
    return TTX_ERR_BAD_ARG;
  }
  ttx::StepArgs a;
  a.A = d_A; a.m = m; a.n = n; a.r = ell; a.s = 0; a.q = q; a.seed = seed; a.stream_id = 0;
  a.precision = precision; a.method = TTX_METHOD_RSVD; a.core = d_U; a.next = d_SVT;
  ttx::StepOut o;
  ttx_status st = ttx::run_step(h, a, o);
  if (st != TTX_OK) return st;
  if (o.nonconv) return ttx::fail(h, TTX_ERR_NO_CONVERGE, "ttx_rsvd: the Jacobi eigensolver did not converge");
  if (d_S) {
    hipError_t e = hipMemcpyAsync(d_S, o.sigma, ell * sizeof(float), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) { h->err = hipGetErrorString(e); return TTX_ERR_HIP; }
  }
  return TTX_OK;
}

ttx_status ttx_permute_231(ttx_handle hh, const float* d_in, float* d_out, int64_t r0, int64_t r1, int64_t n) {
  Handle* h = H(hh);
  if (!h || !d_in || !d_out || r0 <= 0 || r1 <= 0 || n <= 0) return TTX_ERR_BAD_ARG;
  ttx::launch_permute_231(d_in, d_out, r0, r1, n, h->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) { h->err = hipGetErrorString(e); return TTX_ERR_HIP; }
  return TTX_OK;
}

ttx_status ttx_fill_normal(ttx_handle hh, float* d_out, int64_t count, uint64_t seed, uint32_t stream, float scale) {
  Handle* h = H(hh);
  if (!h || !d_out || count < 0) return TTX_ERR_BAD_ARG;
  if (count == 0) return TTX_OK;
  ttx::launch_fill_normal(d_out, count, seed, stream, scale, h->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) { h->err = hipGetErrorString(e); return TTX_ERR_HIP; }
  return TTX_OK;
}

ttx_status ttx_omega(ttx_handle hh, float* d_out, int64_t n, int32_t ell, uint64_t seed, uint32_t stream) {
  Handle* h = H(hh);
  if (!h || !d_out || n <= 0 || ell <= 0) return TTX_ERR_BAD_ARG;
  ttx::launch_omega(d_out, n, ell, seed, stream, h->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) { h->err = hipGetErrorString(e); return TTX_ERR_HIP; }
  return TTX_OK;
}

ttx_status ttx_tt_generate(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks, float std_,
                           float noise_rel, uint64_t seed, float* d_A, float* const* d_cores) {
  if (!h || !dims || !ranks) return TTX_ERR_BAD_ARG;
  if (ranks[0] != 1 || ranks[order] != 1) { H(h)->err = "TT ranks need r_0 = r_d = 1"; return TTX_ERR_BAD_ARG; }
  return ttx::generate(H(h), order, dims, ranks, std_, noise_rel, seed, d_A, d_cores, false);
}

ttx_status ttx_tr_generate(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks, float std_,
                           float noise_rel, uint64_t seed, float* d_A, float* const* d_cores) {
  if (!h || !dims || !ranks) return TTX_ERR_BAD_ARG;
  if (ranks[0] != ranks[order]) { H(h)->err = "TR ranks need r_0 = r_d"; return TTX_ERR_BAD_ARG; }
  return ttx::generate(H(h), order, dims, ranks, std_, noise_rel, seed, d_A, d_cores, true);
}

ttx_status ttx_tt_generate_ref(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks, uint32_t seed,
                               float* d_A, float* const* d_cores) {
  if (!h || !dims || !ranks) return TTX_ERR_BAD_ARG;
  if (ranks[0] != 1 || ranks[order] != 1) { H(h)->err = "TT ranks need r_0 = r_d = 1"; return TTX_ERR_BAD_ARG; }
  return ttx::generate(H(h), order, dims, ranks, 1.0f, 0.0f, seed, d_A, d_cores, false, 1);
}

ttx_status ttx_tr_generate_ref(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks, uint32_t seed,
                               float* d_A, float* const* d_cores) {
  if (!h || !dims || !ranks) return TTX_ERR_BAD_ARG;
  if (ranks[0] != ranks[order]) { H(h)->err = "TR ranks need r_0 = r_d"; return TTX_ERR_BAD_ARG; }
  return ttx::generate(H(h), order, dims, ranks, 1.0f, 0.0f, seed, d_A, d_cores, true, 1);
}

ttx_status ttx_tt_rel_error(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks,
                            const float* const* d_cores, const float* d_A, double* err) {
  if (!h || !dims || !ranks || !d_cores || !d_A || !err || order < 2 || order > TTX_MAX_ORDER) return TTX_ERR_BAD_ARG;
  return ttx::rel_error(H(h), order, dims, ranks, d_cores, d_A, false, err);
}

ttx_status ttx_tr_rel_error(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks,
                            const float* const* d_cores, const float* d_A, double* err) {
  if (!h || !dims || !ranks || !d_cores || !d_A || !err || order < 2 || order > TTX_MAX_ORDER) return TTX_ERR_BAD_ARG;
  return ttx::rel_error(H(h), order, dims, ranks, d_cores, d_A, true, err);
}

ttx_status ttx_tt_round(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks_in,
                        const float* const* d_cores_in, double eps, const int32_t* max_ranks, float* const* d_cores_out,
                        int32_t* ranks_out, float* h_sigma) {
  if (!h || !dims || !ranks_in || !d_cores_in || !d_cores_out || !ranks_out) return TTX_ERR_BAD_ARG;
  return ttx::tt_round(H(h), order, dims, ranks_in, d_cores_in, eps, max_ranks, d_cores_out, ranks_out, h_sigma);
}

ttx_status ttx_tt_full(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks,
                       const float* const* d_cores, float* d_out) {
  if (!h || !dims || !ranks || !d_cores || !d_out || order < 2 || order > TTX_MAX_ORDER) return TTX_ERR_BAD_ARG;
  return ttx::full(H(h), order, dims, ranks, d_cores, d_out, false);
}

ttx_status ttx_tr_full(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks,
                       const float* const* d_cores, float* d_out) {
  if (!h || !dims || !ranks || !d_cores || !d_out || order < 2 || order > TTX_MAX_ORDER) return TTX_ERR_BAD_ARG;
  return ttx::full(H(h), order, dims, ranks, d_cores, d_out, true);
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// rank model entry points
// ------------------------------------------------------------------------------------------------
extern "C" {

ttx_status ttx_mlp_init(ttx_handle hh, uint64_t seed) {
  Handle* h = H(hh);
  if (!h) return TTX_ERR_BAD_ARG;
  TTX_TRY(ttx::mlp_buffers(h, 1));
  ttx::mlp_init(ttx::P<float>(h->mP), TTX_MLP_IN, TTX_MLP_HIDDEN, TTX_MLP_OUT, seed, h->stream);
  TTX_HIP(h, hipMemsetAsync(h->mM1.p, 0, TTX_MLP_PARAMS * sizeof(float), h->stream));
  TTX_HIP(h, hipMemsetAsync(h->mM2.p, 0, TTX_MLP_PARAMS * sizeof(float), h->stream));
  TTX_TRY(ttx::check_launch(h, "mlp init"));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  h->mlp_t = 0;
  h->mlp_ready = true;
  return TTX_OK;
}

ttx_status ttx_mlp_set_params(ttx_handle hh, const float* hp) {
  Handle* h = H(hh);
  if (!h || !hp) return TTX_ERR_BAD_ARG;
  TTX_TRY(ttx::mlp_buffers(h, 1));
  TTX_HIP(h, hipMemcpyAsync(h->mP.p, hp, TTX_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, h->stream));
  TTX_HIP(h, hipMemsetAsync(h->mM1.p, 0, TTX_MLP_PARAMS * sizeof(float), h->stream));
  TTX_HIP(h, hipMemsetAsync(h->mM2.p, 0, TTX_MLP_PARAMS * sizeof(float), h->stream));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  h->mlp_t = 0;
  h->mlp_ready = true;
  return TTX_OK;
}

ttx_status ttx_mlp_get_params(ttx_handle hh, float* hp) {
  Handle* h = H(hh);
  if (!h || !hp) return TTX_ERR_BAD_ARG;
  if (!h->mlp_ready) return ttx::fail(h, TTX_ERR_BAD_ARG, "no rank model");
  TTX_HIP(h, hipMemcpyAsync(hp, h->mP.p, TTX_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  return TTX_OK;
}

ttx_status ttx_mlp_load(ttx_handle hh, const char* path) {
  Handle* h = H(hh);
  if (!h || !path) return TTX_ERR_BAD_ARG;
  std::vector<float> v(TTX_MLP_PARAMS);
  FILE* f = std::fopen(path, "rb");
  if (!f) return ttx::fail(h, TTX_ERR_BAD_ARG, std::string("cannot open ") + path);
  const size_t got = std::fread(v.data(), sizeof(float), v.size(), f);
  const bool extra = std::fgetc(f) != EOF;
  std::fclose(f);
  if (got != v.size() || extra) return ttx::fail(h, TTX_ERR_BAD_ARG, "rank model file has the wrong size");
  return ttx_mlp_set_params(hh, v.data());
}

static ttx_status mlp_check(Handle* h, const float* X, int32_t batch) {
  if (!h) return TTX_ERR_BAD_ARG;
  if (!X || batch <= 0 || batch > TTX_MLP_MAX_BATCH) return ttx::fail(h, TTX_ERR_BAD_ARG, "bad batch");
  if (!h->mlp_ready) return ttx::fail(h, TTX_ERR_BAD_ARG, "no rank model (ttx_mlp_init / ttx_mlp_load)");
  return ttx::mlp_buffers(h, batch);
}

ttx_status ttx_mlp_forward(ttx_handle hh, const float* X, int32_t batch, float* Y) {
  Handle* h = H(hh);
  TTX_TRY(mlp_check(h, X, batch));
  if (!Y) return ttx::fail(h, TTX_ERR_BAD_ARG, "null output");
  const ttx::MlpAct a = ttx::mlp_act(h, batch);
  ttx::mlp_forward(ttx::P<float>(h->mP), TTX_MLP_IN, TTX_MLP_HIDDEN, TTX_MLP_OUT, X, batch, a.Z1, a.H1, a.Z2, a.H2, Y,
                   h->stream);
  TTX_TRY(ttx::check_launch(h, "mlp forward"));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  return TTX_OK;
}

static ttx_status mlp_grad(Handle* h, const float* X, const float* T, int32_t batch, float* grad, double* loss) {
  const ttx::MlpAct a = ttx::mlp_act(h, batch);
  const float* P = ttx::P<float>(h->mP);
  ttx::mlp_forward(P, TTX_MLP_IN, TTX_MLP_HIDDEN, TTX_MLP_OUT, X, batch, a.Z1, a.H1, a.Z2, a.H2, a.Y, h->stream);
  double* part = ttx::P<double>(h->mPart);
  ttx::mlp_backward(P, TTX_MLP_IN, TTX_MLP_HIDDEN, TTX_MLP_OUT, X, T, batch, a.Z1, a.H1, a.Z2, a.H2, a.Y, a.dY, a.dH,
                    a.dZ, grad, part, h->stream);
  TTX_TRY(ttx::check_launch(h, "mlp backward"));
  if (loss) {
    const int nb = ttx::mlp_loss_blocks((int64_t)batch * TTX_MLP_OUT);
    std::vector<double> hp(nb);
    TTX_HIP(h, hipMemcpyAsync(hp.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    TTX_HIP(h, hipStreamSynchronize(h->stream));
    double s = 0.0;
    for (double v : hp) s += v;
    *loss = s / ((double)batch * TTX_MLP_OUT);
  }
  return TTX_OK;
}

ttx_status ttx_mlp_gradient(ttx_handle hh, const float* X, const float* T, int32_t batch, float* grad, double* loss) {
  Handle* h = H(hh);
  TTX_TRY(mlp_check(h, X, batch));
  if (!T || !grad) return ttx::fail(h, TTX_ERR_BAD_ARG, "null target or gradient");
  TTX_TRY(mlp_grad(h, X, T, batch, grad, loss));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  return TTX_OK;
}

ttx_status ttx_mlp_train_step(ttx_handle hh, const float* X, const float* T, int32_t batch, float lr, double* loss) {
  Handle* h = H(hh);
  TTX_TRY(mlp_check(h, X, batch));
  if (!T || !(lr > 0.0f)) return ttx::fail(h, TTX_ERR_BAD_ARG, "null target or non-positive learning rate");
  TTX_TRY(mlp_grad(h, X, T, batch, ttx::P<float>(h->mG), loss));
  h->mlp_t += 1;
  ttx::mlp_adam(ttx::P<float>(h->mP), ttx::P<float>(h->mG), ttx::P<float>(h->mM1), ttx::P<float>(h->mM2),
                TTX_MLP_PARAMS, lr, h->mlp_t, h->stream);
  return ttx::check_launch(h, "adam");
}

ttx_status ttx_mlp_features(ttx_handle hh, const double* piv, int32_t kp, double rel_tol, int64_t m, int64_t n,
                            float* x) {
  Handle* h = H(hh);
  if (!h || !piv || !x || kp <= 0 || kp > 64 || !(rel_tol > 0.0) || m <= 0 || n <= 0) return TTX_ERR_BAD_ARG;
  ttx::mlp_features(piv, kp, rel_tol, (double)m, (double)n, x, TTX_MLP_IN, h->stream);
  TTX_TRY(ttx::check_launch(h, "features"));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  return TTX_OK;
}

ttx_status ttx_mlp_predict(ttx_handle hh, const double* piv, int32_t kp, double rel_tol, int64_t m, int64_t n,
                           float* y) {
  Handle* h = H(hh);
  if (!h || !piv || !y || kp <= 0 || kp > 64 || !(rel_tol > 0.0) || m <= 0 || n <= 0) return TTX_ERR_BAD_ARG;
  if (!h->mlp_ready) return ttx::fail(h, TTX_ERR_BAD_ARG, "mlp_predict: no parameters (ttx_mlp_init / load)");
  ttx::mlp_predict(ttx::P<float>(h->mP), piv, kp, rel_tol, (double)m, (double)n, y, h->stream);
  TTX_TRY(ttx::check_launch(h, "predict"));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  return TTX_OK;
}

ttx_status ttx_probe(ttx_handle hh, const float* d_A, int64_t m, int64_t n, uint64_t seed, uint32_t stream,
                     double* h_piv) {
  Handle* h = H(hh);
  if (!h || !d_A || !h_piv || m <= 0 || n <= 0) return TTX_ERR_BAD_ARG;
  if (m > ((int64_t)1 << 30)) return ttx::fail(h, TTX_ERR_UNSUPPORTED, "probe: more than 2^30 rows");
  ttx::StepArgs a;
  a.A = d_A;
  a.m = m;
  a.n = n;
  a.seed = seed;
  a.stream_id = stream;
  const ttx::Arith3 ar = ttx::arith_for(TTX_PREC_BF16, 0);
  ttx::StepOut o;
  TTX_TRY(ttx::mlp_buffers(h, 1));
  double* piv = ttx::P<double>(h->mIO);
  TTX_TRY(ttx::probe(h, a, ar, piv, o));
  TTX_HIP(h, hipMemcpyAsync(h_piv, piv, 64 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  TTX_HIP(h, hipStreamSynchronize(h->stream));
  return TTX_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// batched TT-SVD
// ------------------------------------------------------------------------------------------------
namespace ttx {

static ttx_status make_subs(Handle* h, int S) {
  if (!h->bstream && hipStreamCreateWithFlags(&h->bstream, hipStreamNonBlocking) != hipSuccess) {
    h->bstream = nullptr;
    return fail(h, TTX_ERR_HIP, "cannot create the batch stream");
  }
  while ((int)h->subs.size() < S) {
    Handle* s = new Handle();
    s->dev = h->dev;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&s->h_status, 16 * sizeof(int), 0) != hipSuccess ||
        hipHostMalloc((void**)&s->h_sigma, TTX_MAX_ELL * sizeof(float), 0) != hipSuccess ||
        hipHostMalloc((void**)&s->h_lam, TTX_MAX_ELL * sizeof(double), 0) != hipSuccess ||
        hipHostMalloc((void**)&s->h_aux, 16 * sizeof(double), 0) != hipSuccess) {
      (void)hipGetLastError();
      free_handle(s, s->stream != nullptr);
      return fail(h, TTX_ERR_HIP, "cannot create a batch worker stream");
    }
    s->jacobi_max_sweeps = h->jacobi_max_sweeps;
    s->batch_mode = true;
    h->subs.push_back(s);
  }
  return TTX_OK;
}

static uint64_t fnv(uint64_t hsh, const void* p, size_t bytes) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < bytes; ++i) hsh = (hsh ^ c[i]) * 0x100000001B3ull;
  return hsh;
}

// queue every tensor's sweep on its worker stream (no host synchronisation)
static ttx_status batch_queue(Handle* h, int S, const std::vector<ttx_params>& P, const float* const* tensors,
                              float* const* cores, int order) {
  for (size_t i = 0; i < P.size(); ++i) {
    Handle* sub = h->subs[i % S];
    const ttx_status st = sweep(sub, tensors[i], &P[i], cores + i * order, nullptr, nullptr, false);
    if (st != TTX_OK) return fail(h, st, "tensor " + std::to_string(i) + ": " + sub->err);
  }
  return TTX_OK;
}

static ttx_status batch_fork(Handle* h, hipStream_t origin, int S, hipEvent_t fork) {
  TTX_HIP(h, hipEventRecord(fork, origin));
  for (int s = 0; s < S; ++s) TTX_HIP(h, hipStreamWaitEvent(h->subs[s]->stream, fork, 0));
  return TTX_OK;
}
static ttx_status batch_join(Handle* h, hipStream_t origin, int S, std::vector<hipEvent_t>& joins) {
  for (int s = 0; s < S; ++s) {
    TTX_HIP(h, hipEventRecord(joins[s], h->subs[s]->stream));
    TTX_HIP(h, hipStreamWaitEvent(origin, joins[s], 0));
  }
  return TTX_OK;
}

// ------------------------------------------------------------------------------------------------
// grouped batch engine: every tensor of the batch goes through the same sequence of step phases in
// lockstep, and each phase is one launch per kernel variant covering every tensor that has work in
// it (the blocks of all the tensors' plain launches, concatenated).  The tensors' plans are the
// worker sweep's in batch mode (same paths, widths, grids and arguments), so each tensor's cores are
// those of its own sweep, bit for bit.
// ------------------------------------------------------------------------------------------------
enum GKind { G_GRAM_SMALL, G_REDUCE_F32, G_REDUCE_F64, G_SKETCH, G_PROJ, G_GRAM_ROWS, G_CHOL, G_APPLY, G_SYEVJ,
             G_ASSEMBLE, G_NEXT, G_COPY_COLS, G_PROJ_SHORT };

// the launches of one phase, one record per (kernel, template choices, block size), in first-use order
struct GPhase {
  std::vector<GRec> recs;
  template <class J>
  void add(int kind, int ar, int v0, int v1, int v2, int threads, int lds, const J& j) {
    GRec* r = nullptr;
    for (GRec& x : recs)
      if (x.kind == kind && x.ar == ar && x.v0 == v0 && x.v1 == v1 && x.v2 == v2 && x.threads == threads) r = &x;
    if (!r) {
      recs.emplace_back();
      r = &recs.back();
      r->kind = kind; r->ar = ar; r->v0 = v0; r->v1 = v1; r->v2 = v2; r->threads = threads;
      r->jsize = sizeof(J);
    }
    r->lds = std::max(r->lds, lds);
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&j);
    r->jobs.insert(r->jobs.end(), b, b + sizeof(J));
    const int nb = j.gx * j.gy;
    for (int q = 0; q < nb; ++q) r->map.push_back(int2{r->njobs, q});
    r->njobs += 1;
  }
  void flush(std::vector<GRec>& out) {
    for (GRec& r : recs) out.push_back(std::move(r));
    recs.clear();
  }
};

// one step of one tensor in the grouped plan (the worker sweep's choices in batch mode)
struct BStep {
  int path = 0, m = 0, r = 0, k = 0, kp = 32, kpn = 32, kz = 4;
  int64_t n = 0, ggram = 0, gproj = 0, nb = 0;
  SketchPlan sp;
  const float* A = nullptr;
  float* next = nullptr;
  float* core = nullptr;
  uint32_t stream = 0;
};
struct BWork {   // one tensor's workspace in the batch arena
  float *part = nullptr, *Y = nullptr, *Qc = nullptr, *Qn = nullptr, *Z = nullptr, *Ur = nullptr;
  float* nxt[2] = {nullptr, nullptr};
  double *gpart = nullptr, *W1 = nullptr, *W2 = nullptr, *Rinv = nullptr, *G = nullptr, *V = nullptr, *lam = nullptr;
  unsigned long long* keys = nullptr;
  float* sig = nullptr;   // [d - 1][TTX_MAX_ELL]
  int* stat = nullptr;    // [d - 1][4]
};

// the schedule of one tensor (sizes only; pointers are set once the arena exists)
static ttx_status bplan_tensor(Handle* h, const ttx_params* p, std::vector<BStep>& steps, size_t sz[16]) {
  const int d = p->order;
  int64_t total = 1;
  for (int k = 0; k < d; ++k) total *= p->dims[k];
  int64_t n = total;
  int r_prev = 1;
  for (int q = 0; q < 16; ++q) sz[q] = 0;
  auto mx = [&](int q, size_t v) { sz[q] = std::max(sz[q], v); };
  steps.assign(d - 1, BStep());
  for (int k = 0; k < d - 1; ++k) {
    BStep& b = steps[k];
    const int64_t m = (int64_t)r_prev * p->dims[k];
    n /= p->dims[k];
    if (m > (1 << 20)) return fail(h, TTX_ERR_UNSUPPORTED, "batch: unfolding with more than 2^20 rows");
    b.m = (int)m;
    b.n = n;
    const int64_t mn = std::min<int64_t>(m, n);
    b.r = (int)std::min<int64_t>(p->ranks[k + 1], mn);
    int64_t ell = (int64_t)b.r + p->oversampling;
    const bool exact = p->method == TTX_METHOD_EXACT ||
                       (p->method == TTX_METHOD_AUTO && (ell >= mn || mn <= TTX_MAX_ELL));
    if (b.r > TTX_MAX_ELL) return fail(h, TTX_ERR_UNSUPPORTED, "batch: rank above TTX_MAX_ELL");
    b.gproj = proj_grid_batch(m, n);
    if (exact && m <= TTX_MAX_ELL) {
      b.path = TTX_PATH_EXACT_M;
      b.k = b.m;
      b.kpn = kp_for(b.r);
      b.ggram = gram_small_grid_batch(n);
      mx(0, (size_t)b.ggram * m * m * sizeof(double));          // gpart
      mx(1, (size_t)b.kpn * m * sizeof(float));                 // Qn
    } else {
      if (exact && n > TTX_MAX_ELL) return fail(h, TTX_ERR_UNSUPPORTED, "exact path needs min(m, n) <= TTX_MAX_ELL");
      if (exact) {
        b.path = TTX_PATH_EXACT_N;
        b.k = (int)n;
      } else {
        b.path = TTX_PATH_RSVD;
        ell = std::min<int64_t>(std::min<int64_t>(ell, TTX_MAX_ELL), mn);
        b.k = (int)ell;
      }
      b.kp = kp_for(b.k);
      b.kz = (b.k + 3) & ~3;
      b.nb = panel_blocks(m);
      if (b.path == TTX_PATH_RSVD) {
        b.sp = plan_sketch_batch(b.m, n, b.kp, h->batch_cpw);
        mx(2, (size_t)b.sp.slabs * m * b.kp * sizeof(float));    // part
      }
      mx(0, (size_t)b.nb * b.k * b.k * sizeof(double));
      mx(0, (size_t)b.gproj * b.k * b.k * sizeof(double));
      mx(3, (size_t)m * b.kp * sizeof(float));                  // Y
      mx(4, (size_t)b.kp * m * sizeof(float));                  // Qc
      mx(5, (size_t)n * b.kp * sizeof(float));                  // Z
      mx(6, (size_t)m * b.k * sizeof(double));                  // W1, W2
      mx(7, (size_t)b.k * b.k * sizeof(double));                // Rinv
      mx(8, (size_t)b.kz * next_from_z_ld(b.r) * sizeof(float)); // Ur
    }
    mx(9, (size_t)b.k * b.k * sizeof(double));                  // G, V
    if (k < d - 2) mx(10 + (k & 1), (size_t)b.r * n * sizeof(float));   // next[0..1]
    b.stream = h->stream_base + (uint32_t)k;
    r_prev = b.r;
  }
  return TTX_OK;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// the grouped plan of a batch: schedules, workspace arena, phases, upload
static ttx_status gplan_build(Handle* h, const std::vector<ttx_params>& P, const float* const* tensors,
                              float* const* cores, int order, GPlan& G) {
  const int count = (int)P.size();
  const int S = order - 1;
  std::vector<std::vector<BStep>> sch(count);
  std::vector<std::array<size_t, 16>> sz(count);
  size_t total = 0;
  const size_t sigstat = (size_t)count * S * (TTX_MAX_ELL * sizeof(float) + 4 * sizeof(int));
  total += align256(sigstat);
  for (int i = 0; i < count; ++i) {
    size_t z[16];
    TTX_TRY(bplan_tensor(h, &P[i], sch[i], z));
    for (int q = 0; q < 16; ++q) sz[i][q] = z[q];
    // gpart part Y Qc Qn Z W1 W2 Rinv G V lam keys Ur next0 next1
    const size_t fixed = align256(TTX_MAX_ELL * sizeof(double)) + align256(64 * sizeof(unsigned long long));
    total += align256(z[0]) + align256(z[2]) + align256(z[3]) + align256(z[4]) + align256(z[1]) + align256(z[5]) +
             2 * align256(z[6]) + align256(z[7]) + 2 * align256(z[9]) + fixed + align256(z[8]) + align256(z[10]) +
             align256(z[11]);
  }
  TTX_TRY(ensure(h, h->barena, total));
  unsigned char* base = static_cast<unsigned char*>(h->barena.p);
  size_t off = 0;
  auto take = [&](size_t bytes) -> void* {
    void* p = bytes ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  G.sigstat = static_cast<float*>(take(sigstat));
  G.sigstat_bytes = sigstat;
  G.steps = S;
  G.count = count;
  std::vector<BWork> W(count);
  for (int i = 0; i < count; ++i) {
    const auto& z = sz[i];
    BWork& w = W[i];
    w.gpart = static_cast<double*>(take(z[0]));
    w.part = static_cast<float*>(take(z[2]));
    w.Y = static_cast<float*>(take(z[3]));
    w.Qc = static_cast<float*>(take(z[4]));
    w.Qn = static_cast<float*>(take(z[1]));
    w.Z = static_cast<float*>(take(z[5]));
    w.W1 = static_cast<double*>(take(z[6]));
    w.W2 = static_cast<double*>(take(z[6]));
    w.Rinv = static_cast<double*>(take(z[7]));
    w.G = static_cast<double*>(take(z[9]));
    w.V = static_cast<double*>(take(z[9]));
    w.lam = static_cast<double*>(take(TTX_MAX_ELL * sizeof(double)));
    w.keys = static_cast<unsigned long long*>(take(64 * sizeof(unsigned long long)));
    w.Ur = static_cast<float*>(take(z[8]));
    w.nxt[0] = static_cast<float*>(take(z[10]));
    w.nxt[1] = static_cast<float*>(take(z[11]));
    float* ss = G.sigstat + (size_t)i * S * (TTX_MAX_ELL + 4);
    w.sig = ss;
    w.stat = reinterpret_cast<int*>(ss + (size_t)S * TTX_MAX_ELL);
    // operands of each step: the input, then the ping-pong buffers; the last step writes the last core
    const float* cur = tensors[i];
    for (int k = 0; k < S; ++k) {
      BStep& b = sch[i][k];
      b.A = cur;
      b.core = cores[(size_t)i * order + k];
      b.next = (k == S - 1) ? cores[(size_t)i * order + order - 1] : w.nxt[k & 1];
      cur = b.next;
    }
  }
  if (off > total) return fail(h, TTX_ERR_HIP, "internal: batch arena overrun");

  const Arith3 ar = arith_for(P[0].precision, P[0].flags);
  G.recs.clear();
  GPhase ph;
  auto orth_phases = [&](int k, auto sel) {
    // shifted CholeskyQR, CholeskyQR, Newton-Schulz (orth(), one phase per launch of it)
    for (int stage = 0; stage < 6; ++stage) {
      for (int i = 0; i < count; ++i) {
        const BStep& b = sch[i][k];
        if (!sel(i)) continue;
        const BWork& w = W[i];
        int* stat = w.stat + 4 * k;
        const int64_t rpw = ((((int64_t)b.m + b.nb - 1) / b.nb + 63) / 64) * 64;
        const int pb = (int)panel_blocks(b.m);
        const int kc = grp_kcap(b.k);   // LDS capacity class of the panel kernels
        switch (stage) {
          case 0: ph.add(G_GRAM_ROWS, 0, 0, kc, 0, 256, 0, JobGramRows{(int)b.nb, 1, w.Y, b.kp, b.m, b.k, rpw, w.gpart}); break;
          case 1: ph.add(G_CHOL, 0, 0, 0, 0, 256, 0, JobChol{1, 1, w.gpart, b.nb, b.k, b.m, 1, w.Rinv, stat}); break;
          case 2: ph.add(G_APPLY, 0, 0, kc, 0, 256, 0, JobApply{pb, 1, w.Y, b.kp, b.m, b.k, b.k, w.Rinv, w.W1, nullptr, 0,
                                                              w.gpart, nullptr, 0, b.k}); break;
          case 3: ph.add(G_CHOL, 0, 0, 0, 0, 256, 0, JobChol{1, 1, w.gpart, b.nb, b.k, b.m, 0, w.Rinv, stat}); break;
          case 4: ph.add(G_APPLY, 0, 1, kc, 0, 256, 0, JobApply{pb, 1, w.W1, b.k, b.m, b.k, b.k, w.Rinv, w.W2, nullptr, 0,
                                                              w.gpart, nullptr, 0, b.k}); break;
          default: ph.add(G_APPLY, 0, 2, kc, 0, 256, 0, JobApply{pb, 1, w.W2, b.k, b.m, b.k, b.kp, nullptr, nullptr, w.Qc,
                                                               b.m, nullptr, w.gpart, b.nb, b.k}); break;
        }
      }
      ph.flush(G.recs);
    }
  };
  auto is_em = [](const BStep& b) { return b.path == TTX_PATH_EXACT_M; };
  auto is_rs = [](const BStep& b) { return b.path == TTX_PATH_RSVD; };
  auto not_em = [](const BStep& b) { return b.path != TTX_PATH_EXACT_M; };
  for (int k = 0; k < S; ++k) {
    // A: exact Gram / copy of the short side / Omega sketch
    for (int i = 0; i < count; ++i) {
      const BStep& b = sch[i][k];
      const BWork& w = W[i];
      if (b.path == TTX_PATH_EXACT_M) {
        int64_t cpw = (b.n + b.ggram - 1) / b.ggram;
        cpw = ((cpw + 63) / 64) * 64;
        ph.add(G_GRAM_SMALL, ar.accurate, b.m <= 32 ? 1 : 2, 0, 0, 256, 0,
               JobGramSmall{(int)b.ggram, 1, b.A, b.m, b.m, b.n, cpw, w.gpart});
      } else if (b.path == TTX_PATH_EXACT_N) {
        ph.add(G_COPY_COLS, 0, 0, 0, 0, 256, 0, JobCopyCols{(int)copy_cols_grid(b.m, b.kp), 1, b.A, b.m, b.m, b.k, w.Y, b.kp});
      } else {
        const SketchPlan& sp = b.sp;
        ph.add(G_SKETCH, ar.range, b.kp / 32, 1, 1, 64 * sp.RB * sp.CG, sp.lds_bytes,
               JobSketch{sp.grid_x, sp.ysplit, b.A, b.m, b.m, b.n, nullptr, 0, P[i].omega_seed, b.stream, b.k, sp.RB,
                         sp.CG, sp.chunks_per_wg, w.part});
      }
    }
    ph.flush(G.recs);
    // B: reductions
    for (int i = 0; i < count; ++i) {
      const BStep& b = sch[i][k];
      const BWork& w = W[i];
      if (b.path == TTX_PATH_EXACT_M) {
        const int64_t cnt = (int64_t)b.m * b.m;
        ph.add(G_REDUCE_F64, 0, 0, 0, 0, 256, 0, JobReduce{(int)((cnt + 31) / 32), 1, w.gpart, b.ggram, cnt, w.G});
      } else if (b.path == TTX_PATH_RSVD) {
        const int64_t cnt = (int64_t)b.m * b.kp;
        ph.add(G_REDUCE_F32, 0, 0, 0, 0, 256, 0, JobReduce{(int)((cnt + 31) / 32), 1, w.part, b.sp.slabs, cnt, w.Y});
      }
    }
    ph.flush(G.recs);
    orth_phases(k, [&](int i) { return not_em(sch[i][k]); });
    // power iterations
    int qmax = 0;
    for (int i = 0; i < count; ++i)
      if (is_rs(sch[i][k])) qmax = std::max(qmax, P[i].power_iters);
    for (int it = 0; it < qmax; ++it) {
      auto sel = [&](int i) { return is_rs(sch[i][k]) && it < P[i].power_iters; };
      for (int i = 0; i < count; ++i) {
        if (!sel(i)) continue;
        const BStep& b = sch[i][k];
        const BWork& w = W[i];
        ph.add(G_PROJ, ar.range, b.kp / 32, 0, 0, 512, 0,
               JobProj{(int)b.gproj, 1, b.A, b.m, b.m, b.n, w.Qc, b.m, 0, w.Z, b.kp, 0, proj_blocks(b.n), nullptr, 0});
      }
      ph.flush(G.recs);
      for (int i = 0; i < count; ++i) {
        if (!sel(i)) continue;
        const BStep& b = sch[i][k];
        const BWork& w = W[i];
        const SketchPlan& sp = b.sp;
        const int fq = b.kp / (16 * sp.RB);
        ph.add(G_SKETCH, ar.range, b.kp / 32, 0, fq <= 1 ? 1 : (fq == 2 ? 2 : 4), 64 * sp.RB * sp.CG, sp.lds_bytes,
               JobSketch{sp.grid_x, sp.ysplit, b.A, b.m, b.m, b.n, w.Z, b.kp, 0, 0, b.k, sp.RB, sp.CG,
                         sp.chunks_per_wg, w.part});
      }
      ph.flush(G.recs);
      for (int i = 0; i < count; ++i) {
        if (!sel(i)) continue;
        const BStep& b = sch[i][k];
        const BWork& w = W[i];
        const int64_t cnt = (int64_t)b.m * b.kp;
        ph.add(G_REDUCE_F32, 0, 0, 0, 0, 256, 0, JobReduce{(int)((cnt + 31) / 32), 1, w.part, b.sp.slabs, cnt, w.Y});
      }
      ph.flush(G.recs);
      orth_phases(k, sel);
    }
    // E: the tail's projection with its Gram, and its reduction
    for (int i = 0; i < count; ++i) {
      const BStep& b = sch[i][k];
      if (!not_em(b)) continue;
      const BWork& w = W[i];
      ph.add(G_PROJ, ar.tail, b.kp / 32, 1, 0, 512, 0,
             JobProj{(int)b.gproj, 1, b.A, b.m, b.m, b.n, w.Qc, b.m, 0, w.Z, b.kz, b.kz, proj_blocks(b.n), w.gpart, b.k});
    }
    ph.flush(G.recs);
    for (int i = 0; i < count; ++i) {
      const BStep& b = sch[i][k];
      if (!not_em(b)) continue;
      const BWork& w = W[i];
      const int64_t cnt = (int64_t)b.k * b.k;
      ph.add(G_REDUCE_F64, 0, 0, 0, 0, 256, 0, JobReduce{(int)((cnt + 31) / 32), 1, w.gpart, b.gproj, cnt, w.G});
    }
    ph.flush(G.recs);
    // F: eigensolver (every tensor)
    for (int i = 0; i < count; ++i) {
      const BStep& b = sch[i][k];
      const BWork& w = W[i];
      ph.add(G_SYEVJ, 0, 0, 0, 0, 1024, 0,
             JobSyevj{1, 1, w.G, b.k, w.lam, w.V, w.stat + 4 * k, w.keys, syevj_threads(b.k)});
    }
    ph.flush(G.recs);
    // G: core assembly
    for (int i = 0; i < count; ++i) {
      const BStep& b = sch[i][k];
      const BWork& w = W[i];
      const int gb = (int)((b.m + 63) / 64);
      float* sig = w.sig + (size_t)k * TTX_MAX_ELL;
      if (is_em(b))
        ph.add(G_ASSEMBLE, 0, 0, 0, 0, 256, 0,
               JobAssemble{gb, 1, nullptr, 0, b.m, b.m, w.V, w.lam, b.r, b.core, nullptr, w.Qn, b.kpn, sig, w.keys});
      else
        ph.add(G_ASSEMBLE, 0, 0, 0, 0, 256, 0,
               JobAssemble{gb, 1, w.Qc, b.m, b.m, b.k, w.V, w.lam, b.r, b.core, w.Ur, nullptr, 0, sig, w.keys});
    }
    ph.flush(G.recs);
    // H: the next unfolding (exact projection, or the remainder from Z)
    for (int i = 0; i < count; ++i) {
      const BStep& b = sch[i][k];
      const BWork& w = W[i];
      if (is_em(b)) {
        ph.add(G_PROJ_SHORT, ar.accurate, b.kpn / 32, (b.m + 15) / 16, 0, 256, 0,
               JobProj{(int)proj_short_grid(b.n, true), 1, b.A, b.m, b.m, b.n, w.Qn, b.m, 1, b.next, 0, b.r, 0, nullptr,
                       0});
      } else {
        bool cols = false;
        const int64_t g = next_from_z_grid(b.n, b.r, &cols, true);
        ph.add(G_NEXT, 0, 0, 0, 0, 256, 0,
               JobNext{(int)g, 1, w.Z, b.kz, b.n, b.k, w.Ur, b.r, b.next, next_from_z_ld(b.r), cols ? 1 : 0});
      }
    }
    ph.flush(G.recs);
  }
  // upload every descriptor and block map in one copy
  size_t blob = 0;
  for (GRec& r : G.recs) {
    r.joff = blob;
    blob = align256(blob + r.jobs.size());
    r.moff = blob;
    blob = align256(blob + r.map.size() * sizeof(int2));
  }
  TTX_TRY(ensure(h, h->bblob, blob));
  TTX_HIP(h, hipStreamSynchronize(h->stream));   // nothing queued earlier still reads the blob
  std::vector<unsigned char> host(blob, 0);
  for (const GRec& r : G.recs) {
    memcpy(host.data() + r.joff, r.jobs.data(), r.jobs.size());
    memcpy(host.data() + r.moff, r.map.data(), r.map.size() * sizeof(int2));
  }
  TTX_HIP(h, hipMemcpy(h->bblob.p, host.data(), blob, hipMemcpyHostToDevice));
  return TTX_OK;
}

// queue the plan's launches on st
static ttx_status gplan_queue(Handle* h, const GPlan& G, hipStream_t st) {
  launch_zero(G.sigstat, (int64_t)G.sigstat_bytes, st);
  const unsigned char* blob = static_cast<const unsigned char*>(h->bblob.p);
  for (const GRec& r : G.recs) {
    if (r.map.empty()) continue;
    GroupLaunch g;
    g.jobs = blob + r.joff;
    g.map = reinterpret_cast<const int2*>(blob + r.moff);
    g.blocks = (int)r.map.size();
    g.threads = r.threads;
    g.lds = r.lds;
    switch (r.kind) {
      case G_GRAM_SMALL: grp_gram_small(r.ar, r.v0, g, st); break;
      case G_REDUCE_F32: grp_reduce(false, g, st); break;
      case G_REDUCE_F64: grp_reduce(true, g, st); break;
      case G_SKETCH: grp_sketch(r.ar, r.v0, r.v1 != 0, r.v2, g, st); break;
      case G_PROJ: grp_proj(r.ar, r.v0, r.v1 != 0, g, st); break;
      case G_PROJ_SHORT: grp_proj_short(r.ar, r.v0, r.v1, g, st); break;
      case G_GRAM_ROWS: grp_gram_rows(false, r.v1, g, st); break;
      case G_CHOL: grp_chol(g, st); break;
      case G_APPLY: grp_apply(r.v0, r.v1, g, st); break;
      case G_SYEVJ: grp_syevj(g, st, h->jacobi_max_sweeps); break;
      case G_ASSEMBLE: grp_assemble(g, st); break;
      case G_NEXT: grp_next(g, st); break;
      case G_COPY_COLS: grp_copy_cols(g, st); break;
      default: return fail(h, TTX_ERR_HIP, "internal: unknown grouped kernel");
    }
  }
  return check_launch(h, "grouped batch");
}

// run a batch on the grouped engine: build (or reuse) its plan, then queue it eagerly on the handle's
// stream or, with TTX_BATCH_GRAPH, as one launch of the plan captured into a graph
static ttx_status grouped_run(Handle* h, const std::vector<ttx_params>& P, const float* const* tensors,
                              float* const* cores, int order, int bflags, uint64_t sig, ttx_batch_report* rep) {
  const int count = (int)P.size();
  const bool graph = (bflags & TTX_BATCH_GRAPH) != 0;
  if (!h->gplan || h->gplan->sig != sig) {
    h->gplan.reset();
    std::unique_ptr<GPlan> G(new GPlan());
    TTX_TRY(gplan_build(h, P, tensors, cores, order, *G));
    G->sig = sig;
    h->gplan = std::move(G);
  }
  GPlan& G = *h->gplan;
  const bool reused = graph && G.gexec != nullptr;
  if (!h->bstream && hipStreamCreateWithFlags(&h->bstream, hipStreamNonBlocking) != hipSuccess) {
    h->bstream = nullptr;
    return fail(h, TTX_ERR_HIP, "cannot create the batch stream");
  }
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // e0, e1, fork, join
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < 4; ++i)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } evg{ev};
  for (int i = 0; i < 4; ++i)
    TTX_HIP(h, i < 2 ? hipEventCreate(&ev[i]) : hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
  if (graph && !G.gexec) {
    if (hipStreamBeginCapture(h->bstream, hipStreamCaptureModeRelaxed) != hipSuccess) {
      (void)hipGetLastError();
      return fail(h, TTX_ERR_HIP, "batch: begin capture");
    }
    const ttx_status qs = gplan_queue(h, G, h->bstream);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(h->bstream, &g);
    if (qs != TTX_OK || ce != hipSuccess || !g) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return qs != TTX_OK ? qs : fail(h, TTX_ERR_HIP, "batch: end capture");
    }
    G.graph = g;
    if (hipGraphInstantiate(&G.gexec, g, nullptr, nullptr, 0) != hipSuccess) {
      G.gexec = nullptr;
      (void)hipGetLastError();
      return fail(h, TTX_ERR_HIP, "batch: graph instantiation");
    }
  }
  TTX_HIP(h, hipEventRecord(ev[0], h->stream));
  if (graph) {
    // the caller's stream -> the graph on the batch stream -> the caller's stream
    TTX_HIP(h, hipEventRecord(ev[2], h->stream));
    TTX_HIP(h, hipStreamWaitEvent(h->bstream, ev[2], 0));
    TTX_HIP(h, hipGraphLaunch(G.gexec, h->bstream));
    TTX_HIP(h, hipEventRecord(ev[3], h->bstream));
    TTX_HIP(h, hipStreamWaitEvent(h->stream, ev[3], 0));
  } else {
    TTX_TRY(gplan_queue(h, G, h->stream));
  }
  TTX_HIP(h, hipEventRecord(ev[1], h->stream));
  TTX_HIP(h, hipEventSynchronize(ev[1]));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
  // the eigensolvers' convergence flags (status word 2 of every step of every tensor)
  std::vector<float> ss(G.sigstat_bytes / sizeof(float));
  TTX_HIP(h, hipMemcpy(ss.data(), G.sigstat, G.sigstat_bytes, hipMemcpyDeviceToHost));
  ttx_batch_report R;
  memset(&R, 0, sizeof(R));
  R.total_ms = ms;
  R.tensors_per_s = ms > 0 ? count / (ms * 1e-3) : 0.0;
  R.p50_ms = R.p99_ms = -1.0;
  R.streams = 1;
  R.graph = graph ? 1 : 0;
  R.graph_reused = reused ? 1 : 0;
  R.tuned_shapes = (int)h->cpw_tuned.size();
  R.sketch_chunks = (int)h->batch_cpw;
  R.engine = 1;
  R.launches = 0;
  for (const GRec& r : G.recs) R.launches += r.map.empty() ? 0 : (r.kind == G_ASSEMBLE ? 2 : 1);
  if (bflags & TTX_BATCH_PER_TENSOR) {
    // each tensor alone through its worker sweep (the same plan), queued work only
    Handle* sub = h->subs[0];
    sub->capture = true;
    std::vector<double> t(count);
    ttx_status st = TTX_OK;
    for (int i = 0; i < count && st == TTX_OK; ++i) {
      if (hipEventRecord(ev[0], sub->stream) != hipSuccess) { st = fail(h, TTX_ERR_HIP, "batch: event"); break; }
      const ttx_status s2 = sweep(sub, tensors[i], &P[i], cores + (size_t)i * order, nullptr, nullptr, false);
      if (s2 != TTX_OK) { st = fail(h, s2, sub->err); break; }
      if (hipEventRecord(ev[1], sub->stream) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess) {
        st = fail(h, TTX_ERR_HIP, "batch: per-tensor run");
        break;
      }
      float ti = 0.0f;
      (void)hipEventElapsedTime(&ti, ev[0], ev[1]);
      t[i] = ti;
    }
    sub->capture = false;
    TTX_TRY(st);
    std::sort(t.begin(), t.end());
    R.p50_ms = t[(size_t)((count - 1) * 0.50)];
    R.p99_ms = t[(size_t)((count - 1) * 0.99)];
  }
  if (rep) *rep = R;
  const int S = G.steps;
  for (int i = 0; i < count; ++i) {
    const int* stat = reinterpret_cast<const int*>(ss.data() + ((size_t)i * S * (TTX_MAX_ELL + 4) + (size_t)S * TTX_MAX_ELL));
    for (int k = 0; k < S; ++k)
      if (stat[4 * k + 2])
        return fail(h, TTX_ERR_NO_CONVERGE, "batch: tensor " + std::to_string(i) + " step " + std::to_string(k) +
                                                ": the Jacobi eigensolver reached its cap of " +
                                                std::to_string(h->jacobi_max_sweeps) + " sweeps before converging");
  }
  return TTX_OK;
}

static ttx_status batch_run(Handle* h, int count, int order, const int64_t* dims, const int32_t* ranks,
                            const float* const* tensors, float* const* cores, const ttx_params* common, int S,
                            int bflags, ttx_batch_report* rep) {
  if (count <= 0 || order < 2 || order > TTX_MAX_ORDER || !dims || !ranks || !tensors || !cores || !common)
    return fail(h, TTX_ERR_BAD_ARG, "batch: bad arguments");
  if (S < 1 || S > 16) return fail(h, TTX_ERR_BAD_ARG, "batch: 1..16 streams");
  if (common->rank_mode != TTX_RANK_FIXED) return fail(h, TTX_ERR_UNSUPPORTED, "batch: fixed ranks only");
  if (common->flags & TTX_FLAG_TIMING) return fail(h, TTX_ERR_UNSUPPORTED, "batch: no per-kernel timing");
  // the range check reads its residual back to the host and chooses widths and power iterations from
  // it: a host round trip per check, which a captured graph could neither contain nor replay
  if (common->flags & TTX_FLAG_ADAPTIVE)
    return fail(h, TTX_ERR_UNSUPPORTED, "batch: no adaptive range finder (fixed widths and power iterations only)");
  // the grouped engine takes every arithmetic but FP16 (whose operand scales are per-step values the
  // grouped jobs do not carry); FP16 and TTX_BATCH_STREAMS run the worker streams
  const bool grouped = !(bflags & TTX_BATCH_STREAMS) && common->precision != TTX_PREC_FP16;
  if (!grouped && (bflags & TTX_BATCH_GRAPH) && count > TTX_BATCH_GRAPH_MAX)
    return fail(h, TTX_ERR_UNSUPPORTED, "batch: graph capture over worker streams takes at most " +
                                            std::to_string(TTX_BATCH_GRAPH_MAX) +
                                            " tensors; split the batch or use the grouped engine");
  std::vector<ttx_params> P(count);
  for (int i = 0; i < count; ++i) {
    ttx_params_init(&P[i], order, dims + (size_t)i * order, ranks + (size_t)i * (order + 1));
    P[i].oversampling = common->oversampling;
    P[i].power_iters = common->power_iters;
    P[i].omega_seed = common->omega_seed;
    P[i].precision = common->precision;
    P[i].method = common->method;
    P[i].rank_mode = TTX_RANK_FIXED;
    P[i].eps = common->eps;
    P[i].flags = common->flags;
    TTX_TRY(validate(h, &P[i], false));
    if (!tensors[i]) return fail(h, TTX_ERR_BAD_ARG, "batch: null tensor");
    for (int k = 0; k < order; ++k)
      if (!cores[(size_t)i * order + k]) return fail(h, TTX_ERR_BAD_ARG, "batch: null core");
  }
  TTX_TRY(make_subs(h, grouped ? 1 : S));
  const bool graph = (bflags & TTX_BATCH_GRAPH) != 0;
  uint64_t sig = 0xCBF29CE484222325ull;
  sig = fnv(sig, &count, sizeof(count));
  sig = fnv(sig, &order, sizeof(order));
  sig = fnv(sig, &S, sizeof(S));
  sig = fnv(sig, dims, sizeof(int64_t) * count * order);
  sig = fnv(sig, ranks, sizeof(int32_t) * count * (order + 1));
  sig = fnv(sig, tensors, sizeof(void*) * count);
  sig = fnv(sig, cores, sizeof(void*) * count * order);
  sig = fnv(sig, &P[0].oversampling, sizeof(int32_t) * 2);
  sig = fnv(sig, &P[0].omega_seed, sizeof(uint64_t));
  sig = fnv(sig, &P[0].precision, sizeof(int32_t) * 2);
  sig = fnv(sig, &P[0].flags, sizeof(int32_t));
  sig = fnv(sig, &h->jacobi_max_sweeps, sizeof(int));
  // autotuner: the sketch's chunks per workgroup (its tile along the columns, which also sets the
  // partial-sum traffic), chosen for this batch by timing one warm eager grouped run per candidate
  if ((bflags & TTX_BATCH_AUTOTUNE) && common->precision != TTX_PREC_FP16 && !h->cpw_tuned.count(sig)) {
    const int64_t cand[] = {8, 16, 32, 64};
    double best_ms = 1e300;
    int64_t best = h->batch_cpw;
    for (int64_t c : cand) {
      h->batch_cpw = c;
      const uint64_t sc = fnv(sig, &c, sizeof(c));
      ttx_batch_report r;
      TTX_TRY(grouped_run(h, P, tensors, cores, order, 0, sc, &r));   // builds the plan, warms
      TTX_TRY(grouped_run(h, P, tensors, cores, order, 0, sc, &r));
      if (r.total_ms < best_ms) {
        best_ms = r.total_ms;
        best = c;
      }
    }
    h->cpw_tuned[sig] = best;
  }
  {
    auto it = h->cpw_tuned.find(sig);
    if (it != h->cpw_tuned.end()) h->batch_cpw = it->second;
  }
  for (Handle* sub : h->subs) sub->batch_cpw = h->batch_cpw;
  sig = fnv(sig, &h->batch_cpw, sizeof(int64_t));
  if (grouped) return grouped_run(h, P, tensors, cores, order, bflags, sig, rep);
  // the workers' buffers the graph's kernels were captured with: any reallocation since (a larger
  // batch, the per-tensor timing loop) invalidates it
  auto workspace_gen = [&]() {
    uint64_t g = 0xCBF29CE484222325ull;
    for (int s = 0; s < S; ++s) g = fnv(g, &h->subs[s]->wsgen, sizeof(uint64_t));
    return g;
  };
  const bool reuse = graph && h->gexec && h->gsig == sig && h->gws == workspace_gen();

  hipEvent_t fork = nullptr, e0 = nullptr, e1 = nullptr;
  std::vector<hipEvent_t> joins(S, nullptr);
  auto cleanup = [&]() {
    for (Handle* sub : h->subs) { sub->capture = false; sub->strict = false; }
    if (fork) (void)hipEventDestroy(fork);
    for (hipEvent_t e : joins) if (e) (void)hipEventDestroy(e);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  ttx_status st = TTX_OK;
  do {
    if (hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) { st = fail(h, TTX_ERR_HIP, "batch: event creation"); break; }
    for (int s = 0; s < S; ++s)
      if (hipEventCreateWithFlags(&joins[s], hipEventDisableTiming) != hipSuccess) {
        st = fail(h, TTX_ERR_HIP, "batch: event creation");
        break;
      }
    if (st != TTX_OK) break;
    for (Handle* sub : h->subs) sub->capture = true;
    if (!reuse) {
      // eager pass: sizes every worker's workspace
      if ((st = batch_fork(h, h->stream, S, fork)) != TTX_OK) break;
      if ((st = batch_queue(h, S, P, tensors, cores, order)) != TTX_OK) break;
      if ((st = batch_join(h, h->stream, S, joins)) != TTX_OK) break;
      if (hipStreamSynchronize(h->stream) != hipSuccess) { st = fail(h, TTX_ERR_HIP, "batch: eager pass"); break; }
      if (graph) {
        if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
        if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
        for (Handle* sub : h->subs) sub->strict = true;
        if (hipStreamBeginCapture(h->bstream, hipStreamCaptureModeRelaxed) != hipSuccess) {
          (void)hipGetLastError();
          st = fail(h, TTX_ERR_HIP, "batch: begin capture");
          break;
        }
        ttx_status qs = batch_fork(h, h->bstream, S, fork);
        if (qs == TTX_OK) qs = batch_queue(h, S, P, tensors, cores, order);
        if (qs == TTX_OK) qs = batch_join(h, h->bstream, S, joins);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(h->bstream, &g);
        for (Handle* sub : h->subs) sub->strict = false;
        if (qs != TTX_OK || ce != hipSuccess || !g) {
          if (g) (void)hipGraphDestroy(g);
          (void)hipGetLastError();
          st = (qs != TTX_OK) ? qs : fail(h, TTX_ERR_HIP, "batch: end capture");
          break;
        }
        h->graph = g;
        if (hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
          h->gexec = nullptr;
          st = fail(h, TTX_ERR_HIP, "batch: graph instantiation");
          break;
        }
        h->gsig = sig;
        h->gws = workspace_gen();
      }
    }
    // the measured run
    if (hipEventRecord(e0, h->stream) != hipSuccess) { st = fail(h, TTX_ERR_HIP, "batch: event"); break; }
    if (graph) {
      // the caller's stream -> the graph on the batch stream -> the caller's stream
      if (hipEventRecord(fork, h->stream) != hipSuccess || hipStreamWaitEvent(h->bstream, fork, 0) != hipSuccess ||
          hipGraphLaunch(h->gexec, h->bstream) != hipSuccess || hipEventRecord(joins[0], h->bstream) != hipSuccess ||
          hipStreamWaitEvent(h->stream, joins[0], 0) != hipSuccess) {
        st = fail(h, TTX_ERR_HIP, "batch: graph launch");
        break;
      }
    } else {
      if ((st = batch_fork(h, h->stream, S, fork)) != TTX_OK) break;
      if ((st = batch_queue(h, S, P, tensors, cores, order)) != TTX_OK) break;
      if ((st = batch_join(h, h->stream, S, joins)) != TTX_OK) break;
    }
    if (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      st = fail(h, TTX_ERR_HIP, "batch: run");
      break;
    }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ttx_batch_report R;
    memset(&R, 0, sizeof(R));
    R.total_ms = ms;
    R.tensors_per_s = ms > 0 ? count / (ms * 1e-3) : 0.0;
    R.p50_ms = R.p99_ms = -1.0;
    R.streams = S;
    R.graph = graph ? 1 : 0;
    R.graph_reused = reuse ? 1 : 0;
    R.tuned_shapes = (int)h->cpw_tuned.size();
    R.sketch_chunks = (int)h->batch_cpw;
    R.engine = 0;
    R.launches = 0;
    if (bflags & TTX_BATCH_PER_TENSOR) {
      // each tensor alone on one worker stream (queued work only, as in the batch)
      Handle* sub = h->subs[0];
      std::vector<double> t(count);
      for (int i = 0; i < count && st == TTX_OK; ++i) {
        if (hipEventRecord(e0, sub->stream) != hipSuccess) { st = fail(h, TTX_ERR_HIP, "batch: event"); break; }
        const ttx_status s2 = sweep(sub, tensors[i], &P[i], cores + (size_t)i * order, nullptr, nullptr, false);
        if (s2 != TTX_OK) { st = fail(h, s2, sub->err); break; }
        if (hipEventRecord(e1, sub->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
          st = fail(h, TTX_ERR_HIP, "batch: per-tensor run");
          break;
        }
        float ti = 0.0f;
        (void)hipEventElapsedTime(&ti, e0, e1);
        t[i] = ti;
      }
      if (st != TTX_OK) break;
      std::sort(t.begin(), t.end());
      R.p50_ms = t[(size_t)((count - 1) * 0.50)];
      R.p99_ms = t[(size_t)((count - 1) * 0.99)];
    }
    if (rep) *rep = R;
  } while (false);
  cleanup();
  return st;
}

}  // namespace ttx

extern "C" ttx_status ttx_batch_tt_svd(ttx_handle hh, int32_t count, int32_t order, const int64_t* dims,
                                       const int32_t* ranks, const float* const* d_tensors, float* const* d_cores,
                                       const ttx_params* common, int32_t streams, int32_t batch_flags,
                                       ttx_batch_report* rep) {
  Handle* h = H(hh);
  if (!h) return TTX_ERR_BAD_ARG;
  return ttx::batch_run(h, count, order, dims, ranks, d_tensors, d_cores, common, streams, batch_flags, rep);
}
