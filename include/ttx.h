/*
 * ttx -- MI355X-native TT-SVD / TR-SVD engine, C ABI.
 *
 * Every entry point takes plain pointers and sizes.  Tensors are FP32, column-major: element
 * (i1, ..., id) of an N1 x ... x Nd tensor lives at i1 + N1 (i2 + N2 (i3 + ...)).  Device pointers
 * are HIP device memory of the current device; all work is queued on the handle's stream and the
 * call returns after the stream has been synchronised, unless stated otherwise.
 *
 * Core layouts (the reference's):
 *   TT core k (k = 0..d-1) is the (r_k N_{k+1}) x r_{k+1} column-major matrix of the 3-way array
 *   (r_k, N_{k+1}, r_{k+1}) (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:194-263, contraction
 *   order :141-172), with r_0 = r_d = 1.
 *   TR core 0 is N1 x (r_0 r_1) with column a + r_0 b (trdecomposition/high_order_GPUs/
 *   trdec6_singleGPU.cu:865,873); cores 1..d-2 as TT; core d-1 is r_{d-1} x (N_d r_0).
 */
#ifndef TTX_H
#define TTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TTX_MAX_ORDER 16
#define TTX_MAX_ELL 64 /* sketch width and rank ceiling of one step */

typedef enum {
  TTX_OK = 0,
  TTX_ERR_BAD_ARG = 1,
  TTX_ERR_OOM = 2,
  TTX_ERR_NO_CONVERGE = 3,        /* the FP64 Jacobi eigensolver of a step reached its sweep cap (30, or
                                     TTX_JACOBI_MAX_SWEEPS in the environment at ttx_create); the outputs
                                     are written but not trusted */
  TTX_RESIDUAL_FALLBACK_USED = 4, /* success; the adaptive range finder had to add passes */
  TTX_ERR_HIP = 5,
  TTX_ERR_UNSUPPORTED = 6
} ttx_status;

typedef enum { TTX_PREC_FP32 = 0, TTX_PREC_BF16 = 1, TTX_PREC_FP16 = 2 } ttx_precision;
typedef enum { TTX_RANK_FIXED = 0, TTX_RANK_DELTA = 1, TTX_RANK_LEARNED = 2 } ttx_rank_mode;
typedef enum { TTX_METHOD_AUTO = 0, TTX_METHOD_EXACT = 1, TTX_METHOD_RSVD = 2 } ttx_method;
typedef enum { TTX_PATH_EXACT_M = 0, TTX_PATH_EXACT_N = 1, TTX_PATH_RSVD = 2 } ttx_path;

typedef struct ttx_handle_s* ttx_handle;

/* kernel classes of the per-kernel timing in ttx_report */
#define TTX_NUM_KCLASS 8
#define TTX_K_SKETCH 0     /* Y = A X streaming pass */
#define TTX_K_PROJ 1       /* Q^T A streaming pass (power pass, remainder, exact projection) */
#define TTX_K_GRAM 2       /* exact-path A A^T streaming pass */
#define TTX_K_PERMUTE 3    /* TR 231 relayout */
#define TTX_K_TALL 4       /* Gram of Z and A_next = U^T Z^T */
#define TTX_K_SMALL 5      /* FP64 CholeskyQR rounds, eigensolver, assembly, reductions */

typedef struct {
  int32_t order;                    /* d, 2..TTX_MAX_ORDER */
  int64_t dims[TTX_MAX_ORDER];      /* N_1..N_d */
  int32_t ranks[TTX_MAX_ORDER + 1]; /* r_0..r_d: fixed ranks, or ceilings in DELTA / LEARNED mode.
                                       TT: r_0 = r_d = 1.  TR: r_0 = r_d = ring rank. */
  int32_t oversampling;             /* s; sketch width ell = r + s */
  int32_t power_iters;              /* q */
  uint64_t omega_seed;              /* Philox key of the Gaussian sketch */
  int32_t precision;                /* ttx_precision */
  int32_t method;                   /* ttx_method: AUTO takes the exact Gram path when ell >= min(m, n) */
  int32_t rank_mode;                /* ttx_rank_mode */
  double eps;                       /* relative accuracy of DELTA / LEARNED ranks: delta = eps / sqrt(d-1) */
  int32_t flags;                    /* TTX_FLAG_* */
} ttx_params;

#define TTX_FLAG_FAST_RANGE 1 /* BF16 mode: single bf16 MFMA per product in the range finder (no hi/lo split) */
#define TTX_FLAG_TIMING 2     /* record per-step times into the report */
/* a-posteriori range check (randomized steps): the sketch carries up to 8 extra Gaussian test columns
 * Y_t = A Omega_t that the basis never sees; mean ||(I - Q Q^T) Y_t||^2 / mean ||Y_t||^2 estimates
 * ||(I - Q Q^T) A||_F^2 / ||A||_F^2 without bias.  Power iterations run only while the estimate is
 * above the step's tolerance, and power_iters becomes their ceiling.  In DELTA rank mode the basis
 * also grows (ell = 16, 32, ..., r + s) while the check fails. */
#define TTX_FLAG_ADAPTIVE 4
/* ttx_tt_svd_host: read the host tensor twice in slices (Gram, then projection) even when a device
 * copy of it would fit; by default one read fills a device copy and the projection reads that */
#define TTX_FLAG_STREAM_TWICE 8
/* BF16 mode: the randomized steps' projection Z = A^T Q (and its fused Gram), and the Gram that the
 * first pass of a sweep takes for a full-rank 32-row first step, with three bf16 products per term
 * (hi hi, lo hi, hi lo) instead of four: the dropped lo lo term is below 2^-16 of the product, so a
 * term carries about 2^-16 relative error instead of 2^-17; a quarter fewer MFMAs in the projection,
 * a tenth fewer in the first pass.  Other exact steps and the remainders keep four products. */
#define TTX_FLAG_TAIL_X3 16
/* decompose this one tensor exactly as ttx_batch_tt_svd decomposes each of a batch (batch grids, the
 * register-buffered sketch, no deferred full-rank core): its cores are then bit for bit the batch's */
#define TTX_FLAG_BATCH_PLAN 32

typedef struct {
  int32_t ranks[TTX_MAX_ORDER + 1];   /* ranks used */
  int32_t ell[TTX_MAX_ORDER];         /* sketch width per step (0 on exact paths) */
  int32_t path[TTX_MAX_ORDER];        /* ttx_path per step */
  int32_t power_iters[TTX_MAX_ORDER]; /* power iterations run per step */
  int32_t zero_pivots[TTX_MAX_ORDER]; /* CholeskyQR pivots that were not positive */
  int32_t jacobi_sweeps[TTX_MAX_ORDER];
  double norm2_in[TTX_MAX_ORDER];     /* ||A_k||_F^2 of the unfolding entering step k (when known, else -1) */
  double step_ms[TTX_MAX_ORDER];      /* with TTX_FLAG_TIMING */
  double total_ms;                    /* with TTX_FLAG_TIMING: whole call on the device */
  int64_t bytes_alg;                  /* algorithmic HBM bytes of the passes that ran */
  /* with TTX_FLAG_TIMING, per kernel class (TTX_K_*): summed device time of its launches (hipEvent
   * pairs on the handle's stream), launch count and algorithmic HBM bytes of those launches */
  double kernel_ms[TTX_NUM_KCLASS];
  int32_t kernel_launches[TTX_NUM_KCLASS];
  int64_t kernel_bytes[TTX_NUM_KCLASS];
  /* TTX_FLAG_ADAPTIVE / LEARNED: estimated relative range residual ||(I - Q Q^T) A_k||_F / ||A_k||_F of
   * the basis that was used (-1 when not checked), number of checks, streaming passes over A_k,
   * and 1 where a learned prediction failed its check and the adaptive range finder took over */
  double range_resid[TTX_MAX_ORDER];
  int32_t range_checks[TTX_MAX_ORDER];
  int32_t passes[TTX_MAX_ORDER];
  int32_t fallback[TTX_MAX_ORDER];
  int32_t predicted_rank[TTX_MAX_ORDER];  /* LEARNED: the network's rank for step k (0 otherwise) */
  /* with TTX_FLAG_TIMING, per kernel class: the launch with the most algorithmic bytes (its bytes and
   * device time) */
  int64_t kernel_top_bytes[TTX_NUM_KCLASS];
  double kernel_top_ms[TTX_NUM_KCLASS];
} ttx_report;

/* handle ---------------------------------------------------------------------------------------- */
ttx_status ttx_create(ttx_handle* h, void* hip_stream);
ttx_status ttx_destroy(ttx_handle h);
ttx_status ttx_set_stream(ttx_handle h, void* hip_stream);
const char* ttx_status_string(int32_t status);
const char* ttx_last_error(ttx_handle h);
int32_t ttx_version(void);

/* default parameters: ranks 1, s = 10, q = 2, seed 1234, BF16, AUTO, FIXED */
void ttx_params_init(ttx_params* p, int32_t order, const int64_t* dims, const int32_t* ranks);
/* number of floats of core k with the ranks in p (TT layout, or TR layout when tr != 0) */
int64_t ttx_core_size(const ttx_params* p, int32_t tr, int32_t k);

/* decompositions -------------------------------------------------------------------------------- */
/* Replaces ttdec_half_lu_device_6 (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:799-958)
 * generalised to order d.  d_A: device input (read only).  d_cores[k]: caller-allocated device
 * buffers of ttx_core_size(p, 0, k) floats (with the ceiling ranks).  h_sigma (optional, host):
 * (d-1) x TTX_MAX_ELL singular values per step, descending, zero-filled. */
ttx_status ttx_tt_svd(ttx_handle h, const float* d_A, const ttx_params* p, float* const* d_cores, float* h_sigma,
                      ttx_report* rep);
/* No counterpart in the reference.  ttx_tt_svd for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): every work buffer is a piece
 * of one static device array of TTX_NOALLOC_SCRATCH_BYTES (256 MiB) that each GPU holds from library
 * load until the process ends; no hipMalloc / hipFree is called.  The caller passes the tensor and the
 * core buffers as for ttx_tt_svd, and gets the same results bit for bit.  Sizes whose workspace bound
 * (every buffer request of every step counted separately) exceeds the array are refused with
 * TTX_ERR_UNSUPPORTED before any work is queued; TTX_FLAG_TIMING is refused too.  Every call on a GPU
 * shares that GPU's array and leaves its contents behind: calls that overlap on one GPU (two handles,
 * two threads) race and must be serialised by the caller.  LEARNED ranks use the handle's loaded model. */
#define TTX_NOALLOC_SCRATCH_BYTES (256u << 20)
ttx_status ttx_tt_svd_noalloc(ttx_handle h, const float* d_A, const ttx_params* p, float* const* d_cores,
                              float* h_sigma, ttx_report* rep);
/* Replaces ttdec_half_lu_host_6 (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:656-796, its slice
 * loops :540-652): the tensor stays in host memory h_A (column-major; pinned, e.g. hipHostMalloc, for
 * copies that overlap the kernels).  The first step (exact Gram on the N_1 side, N_1 <= TTX_MAX_ELL)
 * reads it in column slices of slice_cols columns (0: about 64 MB per slice), the copy of the next
 * slice on the handle's copy stream overlapping the kernels of the current one: once, into a device
 * copy that the projection then reads, when that copy fits beside the remainder with 2 GB to spare
 * (and TTX_FLAG_STREAM_TWICE is not set), else twice through two slice buffers.  Its remainder
 * (r_1 x N_2..N_d, N_1 / r_1 times smaller than the tensor) stays on the device for the rest of the
 * sweep, as in ttx_tt_svd.  Order >= 3.  DELTA / LEARNED: the first rank by the delta rule on the
 * exact spectrum, the later steps with the same per-step budget.  Device memory: two slices (or the
 * device copy), the remainder and the sweep's own working set; these buffers stay with the handle
 * for the next call (ttx_destroy frees them). */
ttx_status ttx_tt_svd_host(ttx_handle h, const float* h_A, const ttx_params* p, int64_t slice_cols,
                           float* const* d_cores, float* h_sigma, ttx_report* rep);
/* Replaces ttdec_half_lu_host_6 of trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:668-826: the TR
 * form of ttx_tt_svd_host (first step of rank r_0 r_1 <= TTX_MAX_ELL from the streamed exact Gram, the
 * 231 relayout of its remainder on the device, then the sweep over r_1 N_2, N_3, .., N_d r_0); fixed
 * ranks; TR core layouts as ttx_tr_svd. */
ttx_status ttx_tr_svd_host(ttx_handle h, const float* h_A, const ttx_params* p, int64_t slice_cols,
                           float* const* d_cores, float* h_sigma, ttx_report* rep);
/* Replaces ttdec_half_lu_device_6 of trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:827-1004:
 * first step of rank r_0 r_1, the 231 relayout, then the TT sweep; TR core layouts above.  DELTA /
 * LEARNED ranks (the commented rule of trdecomposition/third_order_GPU/unopt_GPU/
 * trdec_gpu_float_unopt2_rank.cu:92,222-243, delta^2 = eps^2 / (d-1) ||A||_F^2 per step): ranks[1..d-1]
 * are ceilings; the ring rank r_0 = r_d stays as given and the first step keeps r_0 ceil(R / r_0)
 * columns for its delta rank R, so r_1 = ceil(R / r_0); the report holds the ranks used. */
ttx_status ttx_tr_svd(ttx_handle h, const float* d_A, const ttx_params* p, float* const* d_cores, float* h_sigma,
                      ttx_report* rep);
/* Replaces basicRandSVD_LU (ttdecomposition/third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:557-775):
 * ell = k + s; d_U m x ell, d_S ell, d_SVT ell x n (all device, column-major), A ~= U diag(S)... with
 * SVT = diag(S) V^T. */
ttx_status ttx_rsvd(ttx_handle h, const float* d_A, int64_t m, int64_t n, int32_t k, int32_t s, int32_t q,
                    uint64_t seed, int32_t precision, float* d_U, float* d_S, float* d_SVT);
/* One sweep step on an m x n unfolding: core (m x r) and remainder (r x n); r chosen as in p. */
ttx_status ttx_step(ttx_handle h, const float* d_A, int64_t m, int64_t n, int32_t r, int32_t s, int32_t q,
                    uint64_t seed, uint32_t stream_id, int32_t precision, int32_t method, float* d_core,
                    float* d_next, float* h_sigma, int32_t* path_out);

/* relayout and utilities ------------------------------------------------------------------------- */
/* tensorToMode231 (trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:49-58): (r0, r1, n) -> (r1, n, r0) */
ttx_status ttx_permute_231(ttx_handle h, const float* d_in, float* d_out, int64_t r0, int64_t r1, int64_t n);
/* element e = scale * normal4(seed, stream, e / 4, 0)[e % 4] (Philox4x32-10 + Box-Muller) */
ttx_status ttx_fill_normal(ttx_handle h, float* d_out, int64_t count, uint64_t seed, uint32_t stream, float scale);
/* the sketch matrix Omega, n x ell row-major (entry (j, c) at j * ell + c) */
ttx_status ttx_omega(ttx_handle h, float* d_out, int64_t n, int32_t ell, uint64_t seed, uint32_t stream);
/* synthetic tensors: cores N(0, std^2) from streams 0x100 + k (TT) / 0x200 + k (TR), contracted,
 * then A += noise_rel * rms(A) * N(0,1) from stream 0xFFFF.  d_cores (optional) receives the cores. */
ttx_status ttx_tt_generate(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks, float std_,
                           float noise_rel, uint64_t seed, float* d_A, float* const* d_cores);
ttx_status ttx_tr_generate(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks, float std_,
                           float noise_rel, uint64_t seed, float* d_A, float* const* d_cores);
/* the reference's own inputs (generalTTTensor, ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:185-271;
 * generalTRTensor, trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:231-329): cores with entries
 * (float)(rand() / RAND_MAX) * 2 - 1 from glibc's rand() after srand(seed) (the reference seeds 1134),
 * drawn on the host in the reference's order G_{d-1}, G_d, G_{d-2}, ..., G_1 (order 6: G5, G6, G4, G3,
 * G2, G1) and each filled in its storage order (the core layouts above), then contracted on the device
 * in FP32 as ttx_tt_full / ttx_tr_full.  The entries are bit-identical to the reference's; the tensor
 * differs from the reference's only by the FP32 summation order of the contraction. */
ttx_status ttx_tt_generate_ref(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks, uint32_t seed,
                               float* d_A, float* const* d_cores);
ttx_status ttx_tr_generate_ref(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks, uint32_t seed,
                               float* d_A, float* const* d_cores);
/* ||A - reconstruction(cores)||_F / ||A||_F in FP64 without writing the reconstruction (calMSE,
 * ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:134-182, non-destructive) */
ttx_status ttx_tt_rel_error(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks,
                            const float* const* d_cores, const float* d_A, double* err);
ttx_status ttx_tr_rel_error(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks,
                            const float* const* d_cores, const float* d_A, double* err);
/* TT rounding (no counterpart in the reference, whose cores are only freed, ttdec6_singleGPU.cu:947-952;
 * the consumer that recompresses them, e.g. after adding two TTs): Oseledets 2011, TT-rounding
 * (Algorithm 2), right-to-left orthogonalization (shifted CholeskyQR2 of each core's transpose) then
 * left-to-right exact truncation steps with the delta rule, delta^2 = eps^2 ||A||_F^2 / (d - 1), so
 * ||A - round(A)||_F <= eps ||A||_F.  max_ranks (optional, d + 1 entries) caps the ranks.  The input
 * cores are not modified; d_cores_out[k] must hold r_{k-1} N_k r_k floats at the input ranks; ranks_out
 * (d + 1) receives the new ranks, h_sigma (optional, (d - 1) x TTX_MAX_ELL) the singular values of each
 * bond.  TT ranks at most TTX_MAX_ELL, r_{k-1} <= N_k r_k. */
ttx_status ttx_tt_round(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks_in,
                        const float* const* d_cores_in, double eps, const int32_t* max_ranks, float* const* d_cores_out,
                        int32_t* ranks_out, float* h_sigma);
/* contract cores into a full tensor (d_out: prod N floats) */
ttx_status ttx_tt_full(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks,
                       const float* const* d_cores, float* d_out);
ttx_status ttx_tr_full(ttx_handle h, int32_t order, const int64_t* dims, const int32_t* ranks,
                       const float* const* d_cores, float* d_out);

/* batched TT-SVD -----------------------------------------------------------------------------------
 * Many tensors of uneven shapes and ranks (the reference decomposes one tensor per call,
 * ttdec6_singleGPU.cu:799).  Fixed ranks, widths and power iterations only (no host round trips
 * inside): DELTA / LEARNED rank modes, TTX_FLAG_ADAPTIVE and TTX_FLAG_TIMING are refused with
 * TTX_ERR_UNSUPPORTED.
 * Grouped engine (the default; FP32 and BF16): every tensor goes through the same sequence of step
 * phases at once, and each phase is one launch per kernel variant whose blocks are the blocks of all
 * the tensors' own launches (a block map and one job descriptor per tensor, built and uploaded once
 * per batch signature), so a batch is about a hundred launches whatever its size, and each tensor's
 * cores equal those of its own sweep in batch mode bit for bit.  With TTX_BATCH_GRAPH those launches
 * are captured once into a hipGraph and every call with the same tensors, shapes and parameters
 * relaunches it as one launch.  `streams` is not used.  TTX_BATCH_AUTOTUNE (any engine, not FP16):
 * the sketch's column chunks per workgroup (8, 16, 32 or 64) are chosen per batch signature by timing
 * one warm eager grouped run of the batch with each, and kept for the handle's lifetime (ttx_tt_svd
 * with TTX_FLAG_BATCH_PLAN on the same handle then uses the last batch's choice).
 * Worker-stream engine (TTX_BATCH_STREAMS, and every FP16 batch): the tensors' sweeps spread
 * round-robin over `streams` worker streams, each with its own workspace; TTX_BATCH_GRAPH captures
 * them (fork / join) for batches of at most TTX_BATCH_GRAPH_MAX tensors, and a graph is relaunched
 * only while no worker buffer has been reallocated since its capture.
 * Both engines run each tensor's sweep in batch mode: the streaming passes take batch grids (a few
 * large workgroups per tensor), and a full-rank exact step is projected like any other.
 * dims: count x order, ranks: count x (order + 1) (r_0 = r_d = 1), d_cores: count x order device
 * pointers (ttx_core_size of each tensor's parameters). */
#define TTX_BATCH_GRAPH 1
/* largest batch that TTX_BATCH_GRAPH captures on worker streams (the size its tests cover; the launch
 * of one 512-tensor worker-stream graph faulted the GPU, so larger ones are refused with
 * TTX_ERR_UNSUPPORTED; the grouped engine's graph has no such limit) */
#define TTX_BATCH_GRAPH_MAX 16
#define TTX_BATCH_PER_TENSOR 2   /* also time each tensor alone: p50 / p99 */
#define TTX_BATCH_AUTOTUNE 4
#define TTX_BATCH_STREAMS 8      /* the worker-stream engine instead of the grouped one */
typedef struct {
  double total_ms;      /* device time of the batch */
  double tensors_per_s;
  double p50_ms, p99_ms; /* TTX_BATCH_PER_TENSOR: device time of one decomposition alone, else -1 */
  int32_t streams;
  int32_t graph;        /* 1: the batch ran as one graph launch */
  int32_t graph_reused; /* 1: an identical earlier batch's graph was relaunched */
  int32_t tuned_shapes; /* batches in the autotuner's table */
  int32_t engine;       /* 1: grouped launches, 0: worker streams */
  int32_t launches;     /* grouped engine: kernel launches per batch */
  int32_t sketch_chunks; /* column chunks per sketch workgroup (autotuned with TTX_BATCH_AUTOTUNE) */
} ttx_batch_report;
ttx_status ttx_batch_tt_svd(ttx_handle h, int32_t count, int32_t order, const int64_t* dims, const int32_t* ranks,
                            const float* const* d_tensors, float* const* d_cores, const ttx_params* common,
                            int32_t streams, int32_t batch_flags, ttx_batch_report* rep);

/* rank model (TTX_RANK_LEARNED) -------------------------------------------------------------------
 * No counterpart in the reference, whose ranks are fixed inputs (ttdec6_singleGPU.cu:977).  A
 * 2 x 128 GELU network maps 32 log-buckets of a 64-column probe sketch's spectrum (the shifted
 * CholeskyQR pivots of A[:, :n_p] Omega, n_p = max(min(n, 512), n / 16)) plus the relative error
 * budget and log2 m, log2 n to (rank / 64, sketch width / 64).  The predicted width seeds the
 * adaptive range finder, whose a-posteriori check judges it; when the check fails the wider widths
 * and power iterations take over (report: fallback).  Training runs on the GPU through the
 * forward / backward / Adam kernels below. */
#define TTX_MLP_IN 40
#define TTX_MLP_HIDDEN 128
#define TTX_MLP_OUT 2
#define TTX_MLP_PARAMS (TTX_MLP_IN * TTX_MLP_HIDDEN + TTX_MLP_HIDDEN + TTX_MLP_HIDDEN * TTX_MLP_HIDDEN + \
                        TTX_MLP_HIDDEN + TTX_MLP_HIDDEN * TTX_MLP_OUT + TTX_MLP_OUT)
#define TTX_MLP_MAX_BATCH 8192
/* He-normal weights from the Philox stream, zero biases, Adam state reset */
ttx_status ttx_mlp_init(ttx_handle h, uint64_t seed);
/* parameters as TTX_MLP_PARAMS host floats: W1 [IN][H], b1, W2 [H][H], b2, W3 [H][OUT], b3 */
ttx_status ttx_mlp_set_params(ttx_handle h, const float* h_params);
ttx_status ttx_mlp_get_params(ttx_handle h, float* h_params);
/* read TTX_MLP_PARAMS little-endian float32 values from a file */
ttx_status ttx_mlp_load(ttx_handle h, const char* path);
/* Y (batch x OUT) = net(X (batch x IN)), device row-major */
ttx_status ttx_mlp_forward(ttx_handle h, const float* d_X, int32_t batch, float* d_Y);
/* gradient of mean((net(X) - T)^2) w.r.t. the parameters into d_grad (TTX_MLP_PARAMS floats); loss
 * (optional, host) receives the mean squared error */
ttx_status ttx_mlp_gradient(ttx_handle h, const float* d_X, const float* d_T, int32_t batch, float* d_grad,
                            double* loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on that gradient; loss optional (a host sync when given) */
ttx_status ttx_mlp_train_step(ttx_handle h, const float* d_X, const float* d_T, int32_t batch, float lr,
                              double* loss);
/* the model's input features from probe pivots d_pivots[0..kp-1] (device, FP64) */
ttx_status ttx_mlp_features(ttx_handle h, const double* d_pivots, int32_t kp, double rel_tol, int64_t m, int64_t n,
                            float* d_x);
/* features and forward pass of one sample in one launch, as a decomposition step predicts its rank
 * and sketch width: d_y[0..TTX_MLP_OUT-1] (device) equals ttx_mlp_features then ttx_mlp_forward */
ttx_status ttx_mlp_predict(ttx_handle h, const double* d_pivots, int32_t kp, double rel_tol, int64_t m, int64_t n,
                           float* d_y);
/* the probe of one unfolding (device A, m x n column-major): pivots R_ii^2 of the shifted Cholesky of
 * (A[:, :n_p] Omega)^T (A[:, :n_p] Omega), Omega n_p x 64 from (seed, stream), FP32-accurate sketch;
 * h_pivots (host) receives 64 values */
ttx_status ttx_probe(ttx_handle h, const float* d_A, int64_t m, int64_t n, uint64_t seed, uint32_t stream,
                     double* h_pivots);

#ifdef __cplusplus
}
#endif

#endif /* TTX_H */
