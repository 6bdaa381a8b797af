// FP64 small dense kernels of one decomposition step (k = sketch width <= 64):
//   * CholeskyQR rounds on an m x k panel: Gram partials per 64-row block, shifted Cholesky with an
//     explicit inverse of the triangular factor, and the panel update W Rinv fused with the Gram of
//     its result (for the next round);
//   * cyclic Jacobi eigensolver for the projected Gram;
//   * core assembly G = Q V[:, :r] with the sign gauge, and the remainder A_next = U_r^T Z^T.
// These replace the reference's cuSOLVER calls: basicQR (geqrf + orgqr,
// ttdecomposition/third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:85-159), basicEig
// (syevd, ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:334-385) and the descending column
// selection matrixInvertColumn (:274-290).
#include "ttx_device.h"
#include "ttx_kernels.h"

namespace ttx {

constexpr int KMAX = 64;
constexpr int RB64 = 64;   // rows per WG of the panel kernels

// ------------------------------------------------------------------------------------------------
// Gram of a panel block: part[blockIdx.x][a][b] = sum over rows i of the block of X[i][a] X[i][b]
// 4 x 4 register blocks over the upper triangle; rows staged through LDS in chunks of 32.
// ------------------------------------------------------------------------------------------------
template <typename T, int KC = KMAX>
__device__ __forceinline__ void block_gram(const double (*xs)[KC + 1], int nrows, int k, double* __restrict__ out) {
  const int kb = (k + 3) / 4;
  const int nb = kb * (kb + 1) / 2;
  for (int bi = threadIdx.x; bi < nb; bi += blockDim.x) {
    int a = 0, rem = bi;
    while (rem >= kb - a) { rem -= kb - a; ++a; }
    const int ca = 4 * a, cb = 4 * (a + rem);
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0;
    // unrolled so the LDS loads of the next rows are in flight during this row's FMAs
#pragma unroll 4
    for (int rr = 0; rr < nrows; ++rr) {
      double xa[4], xb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) { xa[u] = xs[rr][ca + u]; xb[u] = xs[rr][cb + u]; }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[4 * u + v] += xa[u] * xb[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = ca + u, j = cb + v;
        if (i < k && j < k) {
          out[i * k + j] = acc[4 * u + v];
          out[j * k + i] = acc[4 * u + v];
        }
      }
  }
}

// KC: the widest k the LDS is sized for (the plain launch takes KMAX; a grouped launch of narrow jobs
// a smaller KC, so more of its workgroups share a CU)
template <typename T, int KC = KMAX>
__device__ __forceinline__ void gram_rows_body(const T* __restrict__ X, int64_t ldx, int64_t rows, int k,
                                               int64_t rows_per_wg, double* __restrict__ part, const Blk B) {
  __shared__ double xs[RB64][KC + 1];
  const int64_t r0 = (int64_t)B.x * rows_per_wg;
  const int64_t r1 = (r0 + rows_per_wg < rows) ? r0 + rows_per_wg : rows;
  const int kb4 = ((k + 3) / 4) * 4;
  double* out = part + (int64_t)B.x * k * k;
  // accumulate chunk Grams through a small LDS buffer
  __shared__ double gacc[KC * KC];
  for (int e = threadIdx.x; e < k * k; e += 256) gacc[e] = 0.0;
  __shared__ double gchunk[KC * KC];
  for (int64_t rc = r0; rc < r1; rc += RB64) {
    const int nr = (int)((r1 - rc) < RB64 ? (r1 - rc) : RB64);
    __syncthreads();
    for (int e = threadIdx.x; e < RB64 * kb4; e += 256) {
      const int rr = e / kb4, c = e % kb4;
      xs[rr][c] = (rr < nr && c < k) ? (double)X[(rc + rr) * ldx + c] : 0.0;
    }
    __syncthreads();
    block_gram<T, KC>(xs, nr, k, gchunk);
    __syncthreads();
    for (int e = threadIdx.x; e < k * k; e += 256) gacc[e] += gchunk[e];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < k * k; e += 256) out[e] = gacc[e];
}
template <typename T>
__global__ __launch_bounds__(256) void k_gram_rows(const T* __restrict__ X, int64_t ldx, int64_t rows, int k,
                                                   int64_t rows_per_wg, double* __restrict__ part) {
  gram_rows_body<T>(X, ldx, rows, k, rows_per_wg, part, hw_blk());
}
template <typename T, int KC>
__global__ __launch_bounds__(256) void k_gram_rows_grp(const JobGramRows* __restrict__ jobs, const int2* __restrict__ map) {
  Blk B;
  const JobGramRows& j = grp_job(jobs, map, B);
  gram_rows_body<T, KC>(static_cast<const T*>(j.X), j.ldx, j.rows, j.k, j.rpw, j.part, B);
}

int64_t gram_tall_grid(int64_t rows) {
  int64_t g = (rows + 255) / 256;
  const int64_t cap = (int64_t)num_cus() * 2;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return g;
}

void launch_gram_tall_f32(const float* X, int64_t ldx, int64_t rows, int k, double* part, int64_t grid,
                          hipStream_t st) {
  int64_t rpw = (rows + grid - 1) / grid;
  rpw = ((rpw + RB64 - 1) / RB64) * RB64;
  hipLaunchKernelGGL(k_gram_rows<float>, dim3((unsigned)grid), dim3(256), 0, st, X, ldx, rows, k, rpw, part);
}
void launch_gram_tall_f64(const double* X, int64_t ldx, int64_t rows, int k, double* part, int64_t grid,
                          hipStream_t st) {
  int64_t rpw = (rows + grid - 1) / grid;
  rpw = ((rpw + RB64 - 1) / RB64) * RB64;
  hipLaunchKernelGGL(k_gram_rows<double>, dim3((unsigned)grid), dim3(256), 0, st, X, ldx, rows, k, rpw, part);
}


// dst[i][j] = sum_q part[q][i * k + j]: slabs are read whole (coalesced), two at a time, so each
// thread keeps several independent loads in flight
template <int KC = KMAX>
__device__ __forceinline__ void sum_slabs(const double* __restrict__ part, int64_t S, int k,
                                          double (*dst)[KC + 1]) {
  const int kk = k * k;
  double acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = 0.0;
  int64_t q = 0;
  for (; q + 1 < S; q += 2) {
    const double* p0 = part + q * kk;
    const double* p1 = p0 + kk;
    double x0[16], x1[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int e = threadIdx.x + t * (int)blockDim.x;
      x0[t] = (e < kk) ? p0[e] : 0.0;
      x1[t] = (e < kk) ? p1[e] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 16; ++t) acc[t] += x0[t] + x1[t];
  }
  if (q < S) {
    const double* p0 = part + q * kk;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int e = threadIdx.x + t * (int)blockDim.x;
      if (e < kk) acc[t] += p0[e];
    }
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int e = threadIdx.x + t * (int)blockDim.x;
    if (e < kk) dst[e / k][e - (e / k) * k] = acc[t];
  }
}

// 1 / sqrt(d) for a normal d > 0 to FP64 accuracy: the hardware estimate r0 (relative error below
// 1e-7), then r0 (1 - e)^(-1/2) with e = 1 - d r0^2 to third order in e (truncation ~ e^4 / 4);
// measured over 2^20 values across the exponent range: at most 1.24 ulp from 1 / sqrt(d)
__device__ __forceinline__ double rsqrt_f64(double d) {
  const double r0 = __builtin_amdgcn_rsq(d);
  const double e = fma(-(d * r0), r0, 1.0);
  return fma(r0 * e, fma(fma(e, 0.3125, 0.375), e, 0.5), r0);
}

// ------------------------------------------------------------------------------------------------
// sum S Gram partials, (shifted) Cholesky G + s I = R^T R, explicit Rinv = R^-1 (upper, row-major)
//
// One wavefront factors the matrix in registers: lane x holds column x in W[0..63] (the Gram is
// symmetric, so column x is also row x).  Step i (unrolled, so every register index is a
// constant):
//   * pivot d = W[i] of lane i (read across lanes), inv = 1 / sqrt(d + s), computed one step ahead
//     from lane i's own update;
//   * R[i][x] = W[i] inv goes to LDS, and the trailing update W[l] -= R[i][l] R[i][x] (l > i) reads
//     R[i][l] back as an LDS broadcast;
//   * slot i is then free and takes row i of B = L^-1 (L = R^T):
//     B[i][x] = (delta_ix - sum_{j<i} R[j][i] B[j][x]) inv, from the rows of B already in slots j < i.
// The trailing update keeps the full symmetric column, so no lane needs a mask; lanes x <= i carry
// stale values in slots l > i that are overwritten by B rows before they are read.  No barrier
// inside the factorisation; the other waves only sum the partials and write the result.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xFFFFFFFFll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// KP: k rounded up to a multiple of 8; the matrix is padded with the identity to KP, which leaves
// the leading k x k factor unchanged and makes every loop a compile-time one.  INV = false (Rinv
// null) factors only, for callers that want the pivots.
// a: KMAX x (KMAX + 1) LDS doubles (summed Gram, then B, row c, column x, for the output); Rs: at least
// KP x (KMAX + 2) LDS doubles (rows of R as they are produced)
template <int KP, bool INV>
__device__ __forceinline__ void chol_body(const double* __restrict__ part, int64_t S, int k, int64_t m,
                                          int shift_mode, double* __restrict__ Rinv, int* __restrict__ status,
                                          double* __restrict__ pivots, int kinv, double (*a)[KMAX + 1],
                                          double (*Rs)[KMAX + 2]) {
  const int tid = threadIdx.x;
  sum_slabs(part, S, k, a);
  __syncthreads();
  if (tid < 64) {
    const int x = tid;
    double tr = x < k ? a[x][x] : 0.0, md = tr;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      tr += __shfl_xor(tr, o);
      md = fmax(md, __shfl_xor(md, o));
    }
    // shifted CholeskyQR (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa 2020):
    // s = 11 (m k + k (k+1)) u ||X||_2^2 with ||X||_2^2 <= trace(X^T X)
    const double sh = shift_mode ? 11.0 * ((double)m * k + (double)k * (k + 1)) * 1.1102230246251565e-16 * tr : 0.0;
    const double tiny = (md + sh) * 1e-28;
    double W[KP];
#pragma unroll
    for (int l = 0; l < KP; ++l) {
      const double g = a[l][x];   // in bounds for every lane: a is 64 x 65
      W[l] = (x < k && l < k) ? g : (l == x ? 1.0 : 0.0);
    }
    int nzero = 0;
    // pivot of step i + 1, computed during step i: lane i + 1 updates its own diagonal with its own
    // R[i][i + 1] (the same fma as its trailing update), so the next inverse square root does not
    // wait for the LDS round trip of row i
    auto pivot = [&](int i, double dv, double& inv) {
      const bool real = i < k;
      const double d = lane_value(dv, i) + (real ? sh : 0.0);
      const bool ok = d > tiny;
      inv = ok ? rsqrt_f64(d) : 0.0;
      nzero += (ok || !real) ? 0 : 1;
      if (pivots && x == 0 && real) pivots[i] = ok ? d : 0.0;
    };
    double inv;
    pivot(0, W[0], inv);
#pragma unroll
    for (int i = 0; i < KP; ++i) {
      const double r = W[i] * inv;
      Rs[i][x] = r;
      double inv_next = 0.0;
      if (i + 1 < KP) pivot(i + 1, fma(-r, r, W[i + 1]), inv_next);
      // row i of B from the rows j < i (independent of this step's pivot until the last multiply)
      double acc0 = (x == i) ? 1.0 : 0.0, acc1 = 0.0;
      // rows i >= kinv of B are not wanted (only the leading kinv x kinv block of Rinv is)
      if (INV && i < kinv) {
#pragma unroll
        for (int j = 0; j < i; j += 2) {
          acc0 -= Rs[j][i] * W[j];
          if (j + 1 < i) acc1 -= Rs[j + 1][i] * W[j + 1];
        }
      }
      // trailing update of the columns still to be factored
#pragma unroll
      for (int l = i + 1; l < KP; ++l) W[l] = fma(-Rs[i][l], r, W[l]);
      W[i] = (acc0 + acc1) * inv;
      // materialise this step's results here: left alone, the compiler defers each trailing
      // update to the step that reads the slot, holding O(k^2) loaded factors live
#pragma unroll
      for (int l = i; l < KP; ++l) asm volatile("" : "+v"(W[l]));
      inv = inv_next;
    }
    if (INV && x < k) {
#pragma unroll
      for (int c = 0; c < KP; ++c)
        if (c < k) a[c][x] = W[c];
    }
    if (x == 0 && status) status[0] += nzero;
  }
  if (!INV) return;
  __syncthreads();
  // Rinv[i][c] = B[c][i] (upper triangle)
  for (int e = tid; e < k * k; e += 256) {
    const int i = e / k, c = e - (e / k) * k;
    Rinv[e] = (c >= i) ? a[c][i] : 0.0;
  }
}
template <int KP, bool INV>
__global__ __launch_bounds__(256) void k_chol_rinv(const double* __restrict__ part, int64_t S, int k, int64_t m,
                                                   int shift_mode, double* __restrict__ Rinv,
                                                   int* __restrict__ status, double* __restrict__ pivots, int kinv) {
  __shared__ double a[KMAX][KMAX + 1];
  __shared__ double Rs[KP][KMAX + 2];
  chol_body<KP, INV>(part, S, k, m, shift_mode, Rinv, status, pivots, kinv, a, Rs);
}
// every job's factor size in one launch: the body for KP = k rounded up to 8 (the LDS sized for 64)
__global__ __launch_bounds__(256) void k_chol_grp(const JobChol* __restrict__ jobs, const int2* __restrict__ map) {
  __shared__ double a[KMAX][KMAX + 1];
  __shared__ double Rs[KMAX][KMAX + 2];
  Blk B;
  const JobChol& j = grp_job(jobs, map, B);
  switch ((j.k + 7) / 8) {
#define TTX_CHOL(q) case q: chol_body<8 * q, true>(j.part, j.S, j.k, j.m, j.shift, j.Rinv, j.status, nullptr, j.k, a, Rs); break;
    TTX_CHOL(1) TTX_CHOL(2) TTX_CHOL(3) TTX_CHOL(4) TTX_CHOL(5) TTX_CHOL(6) TTX_CHOL(7) TTX_CHOL(8)
#undef TTX_CHOL
    default: break;
  }
}

void launch_chol_inv(const double* part, int64_t S, int k, int64_t m, int shift_mode, double* Rinv, int* status,
                     hipStream_t st, double* pivots, int kinv) {
  if (kinv <= 0 || kinv > k) kinv = k;
  switch ((k + 7) / 8) {
#define TTX_CHOL(q) case q:                                                                                  \
    if (Rinv) hipLaunchKernelGGL((k_chol_rinv<8 * q, true>), dim3(1), dim3(256), 0, st, part, S, k, m, shift_mode,   \
                                 Rinv, status, pivots, kinv);                                                     \
    else hipLaunchKernelGGL((k_chol_rinv<8 * q, false>), dim3(1), dim3(256), 0, st, part, S, k, m, shift_mode,       \
                            Rinv, status, pivots, kinv);                                                          \
    break;
    TTX_CHOL(1) TTX_CHOL(2) TTX_CHOL(3) TTX_CHOL(4) TTX_CHOL(5) TTX_CHOL(6) TTX_CHOL(7) TTX_CHOL(8)
#undef TTX_CHOL
    default: break;
  }
}

// ------------------------------------------------------------------------------------------------
// W_out = W_in Rinv for a 64-row block, plus the Gram partial of W_out (next CholeskyQR round)
// OUT 0: double row-major W_out (stride k); OUT 1: float column-major Qc[c * ldq + i], c < kp
// ------------------------------------------------------------------------------------------------
// KC: the widest k the LDS is sized for (as gram_rows_body)
template <typename TI, int OUT, int KC = KMAX>
__device__ __forceinline__ void apply_body(const TI* __restrict__ Win, int64_t ldi, int64_t rows, int k, int kp,
                                           const double* __restrict__ Rinv, double* __restrict__ Wd,
                                           float* __restrict__ Qc, int64_t ldq, double* __restrict__ part,
                                           const double* __restrict__ gpart, int64_t S, int ldr, const Blk B) {
  __shared__ double ri[KC][KC + 1];
  __shared__ double ws[RB64][KC + 1];
  __shared__ double os[RB64][KC + 1];
  const int tid = threadIdx.x;
  if (gpart) {
    // Newton-Schulz step towards the orthogonal polar factor: M = (3 I - W^T W) / 2 (full matrix)
    sum_slabs<KC>(gpart, S, k, ri);
    __syncthreads();
    for (int e = tid; e < k * k; e += 256) {
      const int i = e / k, j = e - (e / k) * k;
      ri[i][j] = ((i == j) ? 1.5 : 0.0) - 0.5 * ri[i][j];
    }
  } else {
    // Rinv row-major with row stride ldr (>= k): the leading k x k block of a wider factor's inverse
    // is the inverse of its leading block
    for (int e = tid; e < k * k; e += 256) ri[e / k][e - (e / k) * k] = Rinv[(int64_t)(e / k) * ldr + (e - (e / k) * k)];
  }
  const int64_t i0 = (int64_t)B.x * RB64;
  const int nr = (int)((rows - i0) < RB64 ? (rows - i0) : RB64);
  for (int e = tid; e < RB64 * k; e += 256) {
    const int rr = e / k, c = e - (e / k) * k;
    ws[rr][c] = (rr < nr) ? (double)Win[(i0 + rr) * ldi + c] : 0.0;
  }
  __syncthreads();
  // os = ws ri: 4 x 4 output tiles, one per thread (16 row groups x ceil(k / 4) column groups <= 256),
  // sixteen independent FMA chains; a triangular Rinv is zero below the diagonal, so the terms
  // a2 > c of a tile add zeros (the sum over a2 <= c is unchanged)
  {
    const int ncb = (k + 3) / 4;
    if (tid < 16 * ncb) {
      const int r0 = 4 * (tid / ncb), c0 = 4 * (tid % ncb);
      const int top = gpart ? k - 1 : ((c0 + 3 < k - 1) ? c0 + 3 : k - 1);
      double acc[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
#pragma unroll 2
      for (int a2 = 0; a2 <= top; ++a2) {
        double x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { x[u] = ws[r0 + u][a2]; y[u] = ri[a2][c0 + u]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u][v] += x[u] * y[v];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (c0 + v < k) os[r0 + u][c0 + v] = (r0 + u < nr) ? acc[u][v] : 0.0;
    }
  }
  __syncthreads();
  if (OUT == 0) {
    for (int e = tid; e < nr * k; e += 256) Wd[(i0 + e / k) * k + e % k] = os[e / k][e % k];
  } else {
    for (int e = tid; e < RB64 * kp; e += 256) {
      const int rr = e % RB64, c = e / RB64;
      if (rr < nr) Qc[(int64_t)c * ldq + i0 + rr] = (c < k) ? (float)os[rr][c] : 0.0f;
    }
  }
  if (part) block_gram<double, KC>(os, nr, k, part + (int64_t)B.x * k * k);
}
template <typename TI, int OUT>
__global__ __launch_bounds__(256) void k_apply_gram(const TI* __restrict__ Win, int64_t ldi, int64_t rows, int k,
                                                    int kp, const double* __restrict__ Rinv,
                                                    double* __restrict__ Wd, float* __restrict__ Qc, int64_t ldq,
                                                    double* __restrict__ part, const double* __restrict__ gpart,
                                                    int64_t S, int ldr) {
  apply_body<TI, OUT>(Win, ldi, rows, k, kp, Rinv, Wd, Qc, ldq, part, gpart, S, ldr, hw_blk());
}
template <typename TI, int OUT, int KC>
__global__ __launch_bounds__(256) void k_apply_grp(const JobApply* __restrict__ jobs, const int2* __restrict__ map) {
  Blk B;
  const JobApply& j = grp_job(jobs, map, B);
  apply_body<TI, OUT, KC>(static_cast<const TI*>(j.Win), j.ldi, j.rows, j.k, j.kp, j.Rinv, j.Wd, j.Qc, j.ldq, j.part,
                          j.gpart, j.S, j.ldr, B);
}

void launch_apply_gram_f32(const float* Win, int64_t ldi, int64_t rows, int k, const double* Rinv, double* Wout,
                           double* part, hipStream_t st, int ldr) {
  const int64_t grid = (rows + RB64 - 1) / RB64;
  hipLaunchKernelGGL((k_apply_gram<float, 0>), dim3((unsigned)grid), dim3(256), 0, st, Win, ldi, rows, k, k, Rinv,
                     Wout, (float*)nullptr, (int64_t)0, part, (const double*)nullptr, (int64_t)0, ldr > 0 ? ldr : k);
}
void launch_apply_gram_f64(const double* Win, int64_t rows, int k, const double* Rinv, double* Wout, double* part,
                           hipStream_t st) {
  const int64_t grid = (rows + RB64 - 1) / RB64;
  hipLaunchKernelGGL((k_apply_gram<double, 0>), dim3((unsigned)grid), dim3(256), 0, st, Win, (int64_t)k, rows, k, k,
                     Rinv, Wout, (float*)nullptr, (int64_t)0, part, (const double*)nullptr, (int64_t)0, k);
}
void launch_apply_out(const double* Win, int64_t rows, int k, int kp, const double* Rinv, float* Qc, int64_t ldq,
                      hipStream_t st) {
  const int64_t grid = (rows + RB64 - 1) / RB64;
  hipLaunchKernelGGL((k_apply_gram<double, 1>), dim3((unsigned)grid), dim3(256), 0, st, Win, (int64_t)k, rows, k, kp,
                     Rinv, (double*)nullptr, Qc, ldq, (double*)nullptr, (const double*)nullptr, (int64_t)0, k);
}
void launch_apply_ns_out(const double* Win, int64_t rows, int k, int kp, const double* gpart, int64_t S, float* Qc,
                         int64_t ldq, hipStream_t st) {
  const int64_t grid = (rows + RB64 - 1) / RB64;
  hipLaunchKernelGGL((k_apply_gram<double, 1>), dim3((unsigned)grid), dim3(256), 0, st, Win, (int64_t)k, rows, k, kp,
                     (const double*)nullptr, (double*)nullptr, Qc, ldq, (double*)nullptr, gpart, S, k);
}
int64_t panel_blocks(int64_t rows) { return (rows + RB64 - 1) / RB64; }

// ------------------------------------------------------------------------------------------------
// cyclic Jacobi eigensolver (round-robin ordering), FP64, k <= 64, one workgroup: one thread per
// 2 x 2 block of A and one per block of V (np = k2 / 2 pairs, k2 = k rounded up to even; 2 np^2
// threads up to k = 44, else np^2 threads that rotate both).
// A round rotates the np disjoint pairs of the circle-method schedule (pair u of round t joins
// (t + u) mod (k2 - 1) and (t - u) mod (k2 - 1), pair 0 joins t and k2 - 1), so every thread tracks
// its pairs' indices itself.  Per round: every thread loads its block, the threads of the diagonal
// blocks compute the angles (FP32 angle, (c, s) normalised in FP64 so every applied rotation is
// orthogonal to FP64 accuracy); one barrier; A <- J^T A J and V <- V J on the blocks; a second
// barrier.
// The sweep loop ends when the off-diagonal mass after a sweep is at most 0.5e-24 ||A||_F^2 (the
// last round's blocks tile the matrix, so it comes from registers).  status[1] = sweeps, status[2] = 1
// when the sweep cap (max_sweeps) was reached before that test passed, else 0.
// keys[0..63] (optional) are zeroed for the core assembly that follows.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void jacobi_angle(double app, double aqq, double apq, double& c, double& s) {
  c = 1.0;
  s = 0.0;
  if (apq == 0.0) return;
  // tau = (aqq - app) / (2 apq) to FP32 accuracy (the rotation is re-normalised in FP64 below);
  // v_rcp_f64 keeps the quotient independent of the matrix's scale
  const float tau = (float)((aqq - app) * __builtin_amdgcn_rcp(2.0 * apq));
  float t;
  if (!(fabsf(tau) <= 1e18f)) t = 0.0f;
  else t = (tau >= 0.0f ? 1.0f : -1.0f) * __builtin_amdgcn_rcpf(fabsf(tau) + __builtin_amdgcn_sqrtf(1.0f + tau * tau));
  const float c0 = __builtin_amdgcn_rsqf(1.0f + t * t);
  const double cd = (double)c0, sd = (double)(t * c0);
  // c^2 + s^2 = 1 + d with |d| < 2^-21: two Newton steps for 1/sqrt from 1 leave an error below 1e-26
  const double nrm = cd * cd + sd * sd;
  const double r1 = 1.5 - 0.5 * nrm;
  const double rs = r1 * (1.5 - 0.5 * nrm * r1 * r1);
  c = cd * rs;
  s = sd * rs;
}

// NT: the threads that work (syevj_threads(k), a multiple of 64, at most the block's size); the
// block's other waves only keep the barriers.  When 2 np^2 <= NT, threads [0, np^2) rotate the A
// blocks and threads [np^2, 2 np^2) the V blocks (the V update needs only the column angles, so it
// never delays the next round's angles); otherwise each thread rotates its A and its V block.
// The angle of pair u is computed by the thread of the diagonal block (u, u), from the entries it
// has just loaded: no separate angle threads (which all sat in wave 0, the round's critical path).
__device__ __forceinline__ void syevj_body(const double* __restrict__ G, int k, double* __restrict__ lam,
                                           double* __restrict__ Vout, int* __restrict__ status,
                                           unsigned long long* __restrict__ keys, int max_sweeps, const int NT) {
  __shared__ double a[KMAX][KMAX + 1];
  __shared__ double v[KMAX][KMAX + 1];
  __shared__ double cs[KMAX / 2], sn[KMAX / 2];
  __shared__ double red[16];
  __shared__ int done;
  const int NW = NT >> 6;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (keys && tid < 64) keys[tid] = 0ull;
  const int k2 = k + (k & 1);   // an odd size gets one zero row / column (its rotations are identities)
  const int np = k2 / 2, M = k2 - 1;
  double tot = 0.0;
  for (int e = tid < NT ? tid : k2 * k2; e < k2 * k2; e += NT) {
    const int i = e / k2, j = e % k2;
    // the Gram passes accumulate (i, j) and (j, i) in different orders: use the symmetric part
    const double x = (i < k && j < k) ? 0.5 * (G[i * k + j] + G[j * k + i]) : 0.0;
    a[i][j] = x;
    v[i][j] = (i == j) ? 1.0 : 0.0;
    tot += x * x;
  }
  // ||A||_F^2 is invariant under the rotations: one reduction up front
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
  if (lane == 0) red[wv] = tot;
  __syncthreads();
  tot = 0.0;
  for (int w = 0; w < NW; ++w) tot += red[w];
  const int nb = np * np;
  const bool split = 2 * nb <= NT;
  const bool doA = tid < nb;
  const bool doV = split ? (tid >= nb && tid < 2 * nb) : doA;
  const int bid = (split && tid >= nb) ? tid - nb : tid;
  const int bu = (doA || doV) ? bid / np : 0, bw = (doA || doV) ? bid % np : 0;
  // circle-method pair u at round t: (t, M) for u = 0, else ((t + u) mod M, (t - u) mod M); both
  // ends advance by one (mod M) per round, except the fixed end M of pair 0
  auto wrap = [M](int x) { return x >= M ? x - M : x; };
  int sweeps = 0;
  bool converged = !(tot > 0.0);   // a zero matrix is diagonal
  while (sweeps < max_sweeps && tot > 0.0) {
    ++sweeps;
    double off = 0.0;
    int x1 = bu, y1 = bu == 0 ? M : M - bu, x2 = bw, y2 = bw == 0 ? M : M - bw;
    for (int rnd = 0; rnd < M; ++rnd) {
      const int p1 = x1 < y1 ? x1 : y1, q1 = x1 < y1 ? y1 : x1;
      const int p2 = x2 < y2 ? x2 : y2, q2 = x2 < y2 ? y2 : x2;
      double app = 0.0, apq = 0.0, aqp = 0.0, aqq = 0.0, vpp = 0.0, vpq = 0.0, vqp = 0.0, vqq = 0.0;
      // this round's blocks were written before the previous barrier
      if (doA) {
        app = a[p1][p2]; apq = a[p1][q2]; aqp = a[q1][p2]; aqq = a[q1][q2];
        if (bu == bw) {
          double c, s;
          jacobi_angle(app, aqq, apq, c, s);
          cs[bu] = c;
          sn[bu] = s;
        }
      }
      if (doV) { vpp = v[p1][p2]; vpq = v[p1][q2]; vqp = v[q1][p2]; vqq = v[q1][q2]; }
      __syncthreads();
      if (doA) {
        const double c1 = cs[bu], s1 = sn[bu], c2 = cs[bw], s2 = sn[bw];
        const double rpp = c1 * app - s1 * aqp, rpq = c1 * apq - s1 * aqq;
        const double rqp = s1 * app + c1 * aqp, rqq = s1 * apq + c1 * aqq;
        const double npp = c2 * rpp - s2 * rpq, npq = s2 * rpp + c2 * rpq;
        const double nqp = c2 * rqp - s2 * rqq, nqq = s2 * rqp + c2 * rqq;
        a[p1][p2] = npp;
        a[p1][q2] = npq;
        a[q1][p2] = nqp;
        a[q1][q2] = nqq;
        if (rnd == M - 1) {
          off = npq * npq + nqp * nqp;
          if (bu != bw) off += npp * npp + nqq * nqq;
        }
      }
      if (doV) {
        const double c2 = cs[bw], s2 = sn[bw];
        v[p1][p2] = c2 * vpp - s2 * vpq;
        v[p1][q2] = s2 * vpp + c2 * vpq;
        v[q1][p2] = c2 * vqp - s2 * vqq;
        v[q1][q2] = s2 * vqp + c2 * vqq;
      }
      __syncthreads();
      x1 = wrap(x1 + 1);
      x2 = wrap(x2 + 1);
      if (bu != 0) y1 = wrap(y1 + 1);
      if (bw != 0) y2 = wrap(y2 + 1);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) off += __shfl_xor(off, o);
    if (lane == 0) red[wv] = off;
    __syncthreads();
    if (tid == 0) {
      double sum = 0.0;
      for (int w = 0; w < NW; ++w) sum += red[w];
      done = sum <= 0.5e-24 * tot;
    }
    __syncthreads();
    if (done) { converged = true; break; }
  }
  // sort descending: rank of each eigenvalue (ties broken by index)
  for (int i = tid < NT ? tid : k; i < k; i += NT) {
    const double li = a[i][i];
    int rank = 0;
    for (int j = 0; j < k; ++j) {
      const double lj = a[j][j];
      if (lj > li || (lj == li && j < i)) ++rank;
    }
    lam[rank] = li;
    for (int r = 0; r < k; ++r) Vout[r * k + rank] = v[r][i];
  }
  if (tid == 0 && status) {
    status[1] = sweeps;
    status[2] = converged ? 0 : 1;
  }
}
// ------------------------------------------------------------------------------------------------
// the same eigensolver for a single problem (k_syevj), one barrier per round.  np = k2 / 2 pairs
// (k2 = k rounded up to even).  A round rotates the np disjoint pairs of the circle-method schedule (pair u of round t joins (t + u) mod (k2 - 1) and (t - u) mod (k2 - 1),
// pair 0 joins t and k2 - 1), so every lane tracks its pairs' indices itself.
// Lane (bu, bw) owns the 2 x 2 block (row pair bu, column pair bw) of A and of V.  The lanes of a
// row pair share a wavefront: GP = floor(64 / np) row pairs per wave, lane g np + bw of wave w
// holding row pair w GP + g.  Per round every lane
//   * loads its A block, its V block and the entries (p, p), (q, q), (p, q) of its column pair;
//   * computes its column pair's angle (FP32 angle, (c, s) normalised in FP64 so every applied
//     rotation is orthogonal to FP64 accuracy) and reads its row pair's angle from the lane of the
//     diagonal block (bu, bu) in its own wave, a cross-lane read that needs no barrier;
//   * writes its block of J^T A J into the other copy of A and its block of V J into V;
// then one barrier.  A is double-buffered, so no write of a round can reach an entry that another
// lane still reads in that round.  Every lane that computes the angle of pair u gets the same bits
// (same inputs, same instructions): the rotations are those a single owner per pair would apply.
// The sweep loop ends when the off-diagonal mass after a sweep is at most 0.5e-24 ||A||_F^2 (the
// last round's blocks tile the matrix, so it comes from registers).  status[1] = sweeps, status[2] = 1
// when the sweep cap (max_sweeps) was reached before that test passed, else 0.
// keys[0..63] (optional) are zeroed for the core assembly that follows.
// Same rotations as the two-barrier body above, so the same results to the bit.  Its double-buffered
// A (100 KB of LDS) and 68 VGPRs allow one 1024-thread workgroup per CU, so the grouped launch, which
// runs many problems side by side, keeps the two-barrier body (67 KB: two workgroups per CU).
// ------------------------------------------------------------------------------------------------
// Branch-free (apq = 0 selects the identity at the end), so the loads of the three entries are not
// sunk below a branch and waited for one after the other.
__device__ __forceinline__ void jacobi_angle_sel(double app, double aqq, double apq, double& c, double& s) {
  // tau = (aqq - app) / (2 apq) to FP32 accuracy (the rotation is re-normalised in FP64 below);
  // v_rcp_f64 keeps the quotient independent of the matrix's scale
  const float tau = (float)((aqq - app) * __builtin_amdgcn_rcp(2.0 * apq));
  float t;
  if (!(fabsf(tau) <= 1e18f)) t = 0.0f;
  else t = (tau >= 0.0f ? 1.0f : -1.0f) * __builtin_amdgcn_rcpf(fabsf(tau) + __builtin_amdgcn_sqrtf(1.0f + tau * tau));
  const float c0 = __builtin_amdgcn_rsqf(1.0f + t * t);
  const double cd = (double)c0, sd = (double)(t * c0);
  // c^2 + s^2 = 1 + d with |d| < 2^-21: two Newton steps for 1/sqrt from 1 leave an error below 1e-26
  const double nrm = cd * cd + sd * sd;
  const double r1 = 1.5 - 0.5 * nrm;
  const double rs = r1 * (1.5 - 0.5 * nrm * r1 * r1);
  const bool rot = apq != 0.0;
  c = rot ? cd * rs : 1.0;
  s = rot ? sd * rs : 0.0;
}

// NT: the threads that work (syevj1_threads(k), a multiple of 64, at most the block's size); the
// block's other waves only keep the barriers.
__device__ __forceinline__ void syevj1_body(const double* __restrict__ G, int k, double* __restrict__ lam,
                                           double* __restrict__ Vout, int* __restrict__ status,
                                           unsigned long long* __restrict__ keys, int max_sweeps, const int NT) {
  __shared__ double a[2][KMAX][KMAX + 1];
  __shared__ double v[KMAX][KMAX + 1];
  __shared__ double red[16];
  __shared__ int done;
  const int NW = NT >> 6;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (keys && tid < 64) keys[tid] = 0ull;
  const int k2 = k + (k & 1);   // an odd size gets one zero row / column (its rotations are identities)
  const int np = k2 / 2, M = k2 - 1;
  double tot = 0.0;
  for (int e = tid < NT ? tid : k2 * k2; e < k2 * k2; e += NT) {
    const int i = e / k2, j = e % k2;
    // the Gram passes accumulate (i, j) and (j, i) in different orders: use the symmetric part
    const double x = (i < k && j < k) ? 0.5 * (G[i * k + j] + G[j * k + i]) : 0.0;
    a[0][i][j] = x;
    v[i][j] = (i == j) ? 1.0 : 0.0;
    tot += x * x;
  }
  // ||A||_F^2 is invariant under the rotations: one reduction up front
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
  if (lane == 0) red[wv] = tot;
  __syncthreads();
  tot = 0.0;
  for (int w = 0; w < NW; ++w) tot += red[w];
  const int GP = 64 / np;
  const int g = lane / np, bwl = lane - g * np, bul = wv * GP + g;
  const bool act = tid < NT && g < GP && bul < np;
  // idle lanes follow pair 0's indices (in bounds) and write nothing
  const int bu = act ? bul : 0, bw = act ? bwl : 0;
  const int src = act ? g * np + bu : lane;   // the lane of the diagonal block (bu, bu)
  // circle-method pair u at round t: (t, M) for u = 0, else ((t + u) mod M, (t - u) mod M); both
  // ends advance by one (mod M) per round, except the fixed end M of pair 0
  auto wrap = [M](int x) { return x >= M ? x - M : x; };
  int sweeps = 0, cur = 0;
  bool converged = !(tot > 0.0);   // a zero matrix is diagonal
  while (sweeps < max_sweeps && tot > 0.0) {
    ++sweeps;
    double off = 0.0;
    int x1 = bu, y1 = bu == 0 ? M : M - bu, x2 = bw, y2 = bw == 0 ? M : M - bw;
    for (int rnd = 0; rnd < M; ++rnd) {
      const int p1 = x1 < y1 ? x1 : y1, q1 = x1 < y1 ? y1 : x1;
      const int p2 = x2 < y2 ? x2 : y2, q2 = x2 < y2 ? y2 : x2;
      const double(*A0)[KMAX + 1] = a[cur];
      double(*A1)[KMAX + 1] = a[cur ^ 1];
      // this round's entries were written before the previous barrier
      double dpp = A0[p2][p2], dqq = A0[q2][q2], dpq = A0[p2][q2];
      double app = A0[p1][p2], apq = A0[p1][q2], aqp = A0[q1][p2], aqq = A0[q1][q2];
      // all seven loads in flight together: left alone, the compiler sinks the block's four into the
      // store branch, a second LDS round trip after the angles
      asm volatile("" : "+v"(dpp), "+v"(dqq), "+v"(dpq), "+v"(app), "+v"(apq), "+v"(aqp), "+v"(aqq));
      const double vpp = v[p1][p2], vpq = v[p1][q2], vqp = v[q1][p2], vqq = v[q1][q2];
      double c2, s2;
      jacobi_angle_sel(dpp, dqq, dpq, c2, s2);
      const double c1 = __shfl(c2, src), s1 = __shfl(s2, src);
      if (act) {
        v[p1][p2] = c2 * vpp - s2 * vpq;
        v[p1][q2] = s2 * vpp + c2 * vpq;
        v[q1][p2] = c2 * vqp - s2 * vqq;
        v[q1][q2] = s2 * vqp + c2 * vqq;
        const double rpp = c1 * app - s1 * aqp, rpq = c1 * apq - s1 * aqq;
        const double rqp = s1 * app + c1 * aqp, rqq = s1 * apq + c1 * aqq;
        const double npp = c2 * rpp - s2 * rpq, npq = s2 * rpp + c2 * rpq;
        const double nqp = c2 * rqp - s2 * rqq, nqq = s2 * rqp + c2 * rqq;
        A1[p1][p2] = npp;
        A1[p1][q2] = npq;
        A1[q1][p2] = nqp;
        A1[q1][q2] = nqq;
        if (rnd == M - 1) {
          off = npq * npq + nqp * nqp;
          if (bu != bw) off += npp * npp + nqq * nqq;
        }
      }
      __syncthreads();
      cur ^= 1;
      x1 = wrap(x1 + 1);
      x2 = wrap(x2 + 1);
      if (bu != 0) y1 = wrap(y1 + 1);
      if (bw != 0) y2 = wrap(y2 + 1);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) off += __shfl_xor(off, o);
    if (lane == 0) red[wv] = off;
    __syncthreads();
    if (tid == 0) {
      double sum = 0.0;
      for (int w = 0; w < NW; ++w) sum += red[w];
      done = sum <= 0.5e-24 * tot;
    }
    __syncthreads();
    if (done) { converged = true; break; }
  }
  // sort descending: rank of each eigenvalue (ties broken by index)
  const double(*AF)[KMAX + 1] = a[cur];
  for (int i = tid < NT ? tid : k; i < k; i += NT) {
    const double li = AF[i][i];
    int rank = 0;
    for (int j = 0; j < k; ++j) {
      const double lj = AF[j][j];
      if (lj > li || (lj == li && j < i)) ++rank;
    }
    lam[rank] = li;
    for (int r = 0; r < k; ++r) Vout[r * k + rank] = v[r][i];
  }
  if (tid == 0 && status) {
    status[1] = sweeps;
    status[2] = converged ? 0 : 1;
  }
}
__global__ __launch_bounds__(1024) void k_syevj(const double* __restrict__ G, int k, double* __restrict__ lam,
                                               double* __restrict__ Vout, int* __restrict__ status,
                                               unsigned long long* __restrict__ keys, int max_sweeps) {
  syevj1_body(G, k, lam, Vout, status, keys, max_sweeps, (int)blockDim.x);
}
// one 1024-thread block per job, each job working with its own syevj_threads(k)
__global__ __launch_bounds__(1024) void k_syevj_grp(const JobSyevj* __restrict__ jobs, const int2* __restrict__ map,
                                                   int max_sweeps) {
  Blk B;
  const JobSyevj& j = grp_job(jobs, map, B);
  syevj_body(j.G, j.k, j.lam, j.V, j.status, j.keys, max_sweeps, j.nt);
}

// threads for a k x k problem: one per 2 x 2 block of A and one per block of V when both fit in
// a 1024-thread block, else one per block for both; at least one wavefront (the key clearing)
int syevj_threads(int k) {
  const int np = (k + (k & 1)) / 2;
  const int nb = np * np;
  const int t = ((2 * nb <= 1024 ? 2 * nb : nb) + 63) / 64 * 64;
  return t < 64 ? 64 : t;
}

// threads for a k x k problem: whole waves of floor(64 / np) row pairs of np lanes each (at most
// 16 waves for k = 64)
static int syevj1_threads(int k) {
  const int np = k > 1 ? (k + (k & 1)) / 2 : 1;
  const int gp = 64 / np;
  return 64 * ((np + gp - 1) / gp);
}

void launch_syevj(const double* G, int k, double* lam, double* V, int* status, unsigned long long* keys,
                  hipStream_t st, int max_sweeps) {
  hipLaunchKernelGGL(k_syevj, dim3(1), dim3(syevj1_threads(k)), 0, st, G, k, lam, V, status, keys, max_sweeps);
}

// ------------------------------------------------------------------------------------------------
// core assembly with the sign gauge
// pass 1 (64-row blocks): G = Q V[:, :rk] (or V itself when Q = I), written unflipped; per column the
//   entry of largest |value| (first index on ties) is found with a 64-bit atomicMax on the key
//   (|v| bits << 32) | ((0x7FFFFFFF - i) << 1) | sign(v)
// pass 2 (64-row blocks): flip columns by that entry's sign; write Qn (zero-padded copy), Ur, sigma
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void assemble_g_body(const float* __restrict__ Qc, int64_t ldq, int m, int k,
                                                const double* __restrict__ V, int r, float* __restrict__ Gc,
                                                unsigned long long* __restrict__ keys, const Blk B) {
  __shared__ double vs[KMAX][KMAX + 1];
  __shared__ unsigned long long kmax[KMAX];
  const int tid = threadIdx.x;
  const int rk = r < k ? r : k;
  for (int e = tid; e < k * k; e += 256) vs[e / k][e % k] = V[e];
  if (tid < KMAX) kmax[tid] = 0ull;
  __syncthreads();
  const int64_t i0 = (int64_t)B.x * RB64;
  for (int e = tid; e < RB64 * r; e += 256) {
    const int rr = e % RB64, c = e / RB64;
    const int64_t i = i0 + rr;
    if (i >= m) continue;
    double s = 0.0;
    if (c < rk) {
      if (Qc) {
        for (int a2 = 0; a2 < k; ++a2) s += (double)Qc[(int64_t)a2 * ldq + i] * vs[a2][c];
      } else {
        s = vs[i][c];
      }
    }
    const float g = (float)s;
    Gc[(int64_t)c * m + i] = g;
    if (c < rk) {
      const unsigned long long key = ((unsigned long long)__float_as_uint(fabsf(g)) << 32) |
                                     ((unsigned long long)(0x7FFFFFFFu - (unsigned)i) << 1) | (g < 0.0f ? 1ull : 0ull);
      atomicMax(&kmax[c], key);
    }
  }
  __syncthreads();
  if (tid < rk) atomicMax(&keys[tid], kmax[tid]);
}
__global__ __launch_bounds__(256) void k_assemble_g(const float* __restrict__ Qc, int64_t ldq, int m, int k,
                                                    const double* __restrict__ V, int r, float* __restrict__ Gc,
                                                    unsigned long long* __restrict__ keys) {
  assemble_g_body(Qc, ldq, m, k, V, r, Gc, keys, hw_blk());
}

// row stride of Ur: r rounded up to 8, 16, 32 or 64 (the remainder kernel's register width)
__host__ __device__ constexpr int ur_ld(int r) { return r <= 8 ? 8 : r <= 16 ? 16 : r <= 32 ? 32 : 64; }
int next_from_z_ld(int r) { return ur_ld(r); }

__device__ __forceinline__ void assemble_fix_body(const unsigned long long* __restrict__ keys, int m, int k,
                                                  const double* __restrict__ V, const double* __restrict__ lam,
                                                  int r, float* __restrict__ Gc, float* __restrict__ Ur,
                                                  float* __restrict__ Qn, int kpn, float* __restrict__ sigma,
                                                  const Blk B) {
  __shared__ float sg[KMAX];
  const int tid = threadIdx.x;
  const int rk = r < k ? r : k;
  if (tid < KMAX) sg[tid] = (tid < rk && (keys[tid] & 1ull)) ? -1.0f : 1.0f;
  __syncthreads();
  const int64_t i0 = (int64_t)B.x * RB64;
  for (int e = tid; e < RB64 * r; e += 256) {
    const int rr = e % RB64, c = e / RB64;
    const int64_t i = i0 + rr;
    if (i >= m) continue;
    const float g = Gc[(int64_t)c * m + i] * sg[c];
    Gc[(int64_t)c * m + i] = g;
    if (Qn) Qn[(int64_t)c * m + i] = g;
  }
  if (Qn)
    for (int e = tid; e < RB64 * (kpn - r); e += 256) {
      const int rr = e % RB64, c = r + e / RB64;
      const int64_t i = i0 + rr;
      if (i < m) Qn[(int64_t)c * m + i] = 0.0f;
    }
  if (B.x == 0) {
    if (Ur) {
      // round4(k) x ur_ld(r), zero-padded, as k_next_from_z reads it
      const int ld = ur_ld(r), k4 = (k + 3) & ~3;
      for (int e = tid; e < k4 * ld; e += 256) {
        const int a2 = e / ld, c = e % ld;
        Ur[e] = (a2 < k && c < rk) ? (float)(V[a2 * k + c] * (double)sg[c]) : 0.0f;
      }
    }
    if (sigma)
      for (int i = tid; i < k; i += 256) sigma[i] = (float)sqrt(fmax(lam[i], 0.0));
  }
}
__global__ __launch_bounds__(256) void k_assemble_fix(const unsigned long long* __restrict__ keys, int m, int k,
                                                      const double* __restrict__ V, const double* __restrict__ lam,
                                                      int r, float* __restrict__ Gc, float* __restrict__ Ur,
                                                      float* __restrict__ Qn, int kpn, float* __restrict__ sigma) {
  assemble_fix_body(keys, m, k, V, lam, r, Gc, Ur, Qn, kpn, sigma, hw_blk());
}
__global__ __launch_bounds__(256) void k_assemble_g_grp(const JobAssemble* __restrict__ jobs, const int2* __restrict__ map) {
  Blk B;
  const JobAssemble& j = grp_job(jobs, map, B);
  assemble_g_body(j.Qc, j.ldq, j.m, j.k, j.V, j.r, j.Gc, j.keys, B);
}
__global__ __launch_bounds__(256) void k_assemble_fix_grp(const JobAssemble* __restrict__ jobs,
                                                          const int2* __restrict__ map) {
  Blk B;
  const JobAssemble& j = grp_job(jobs, map, B);
  assemble_fix_body(j.keys, j.m, j.k, j.V, j.lam, j.r, j.Gc, j.Ur, j.Qn, j.kpn, j.sigma, B);
}

void launch_assemble(const float* Qc, int64_t ldq, int m, int k, const double* lam, const double* V, int r, float* Gc,
                     float* Ur, float* Qn, int kpn, float* sigma, unsigned long long* keys, hipStream_t st) {
  const int64_t grid = (m + RB64 - 1) / RB64;
  hipLaunchKernelGGL(k_assemble_g, dim3((unsigned)grid), dim3(256), 0, st, Qc, ldq, m, k, V, r, Gc, keys);
  hipLaunchKernelGGL(k_assemble_fix, dim3((unsigned)grid), dim3(256), 0, st, keys, m, k, V, lam, r, Gc, Ur, Qn, kpn,
                     sigma);
}

// ------------------------------------------------------------------------------------------------
// A_next = Ur^T Z^T  (r x n column-major)
// ------------------------------------------------------------------------------------------------
template <int RM>
__device__ __forceinline__ void next_rows_body(const float* __restrict__ Z, int64_t ldz, int64_t n, int k,
                                               const float* __restrict__ Ur, int r, float* __restrict__ out,
                                               const Blk B) {
  // A_next^T = Z U_r in 32-row tiles on the matrix cores: a wave stages rows j0 .. j0 + 31 of Z (one
  // contiguous run of 32 ldz floats, read with 16-byte loads by consecutive lanes) in its LDS tile,
  // then runs k4 / 2 FP32 MFMAs (32x32x2) per 32 output columns: A = the tile (lane: row l % 32,
  // k = 2 s + l / 32), B = U_r (round4(k) x RM, zero-padded), held in registers for the whole call.
  // A lane writes 32 consecutive floats of one output row per accumulator register.
  constexpr int NT = RM > 32 ? 2 : 1;   // 32-column output tiles
  constexpr int P = KMAX + 1;           // LDS pitch of a staged Z row (odd: the column reads are conflict-free)
  __shared__ float zs[4][32 * P];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int k4 = (k + 3) & ~3, kq = k4 / 4;
  float ub[NT][KMAX / 2];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
#pragma unroll
    for (int st = 0; st < KMAX / 2; ++st) {
      const int kk = 2 * st + (lane >> 5), c = 32 * nt + (lane & 31);
      ub[nt][st] = (kk < k4 && c < RM) ? Ur[kk * RM + c] : 0.0f;
    }
  float* zt = zs[w];
  const int64_t tiles = (n + 31) / 32;
  // a tile's 32 kq float4 pieces, piece e = lane + 64 u at row e / kq, columns 4 (e % kq) .. + 3: all
  // loads of a tile are issued at once, unconditionally (row and piece clamped into the matrix, the
  // pieces outside it zeroed after the load), and the next tile's are in flight while this one is
  // staged, multiplied and written (a load-wait-store loop per piece left one load in flight)
  constexpr int NU = 32 * (KMAX / 4) / 64;
  const int npc = 32 * kq;
  auto load_tile = [&](int64_t j0, f32x4 (&v)[NU]) {
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int e = lane + 64 * u;
      const int ec = e < npc ? e : npc - 1;
      const int rr = ec / kq, q = ec - rr * kq;
      const int64_t row = (j0 + rr < n) ? j0 + rr : n - 1;
      v[u] = *reinterpret_cast<const f32x4*>(Z + row * ldz + 4 * q);
    }
  };
  int64_t t = (int64_t)B.x * 4 + w;
  f32x4 vn[NU];
  if (t < tiles) load_tile(t * 32, vn);
  for (; t < tiles; t += (int64_t)B.gx * 4) {
    const int64_t j0 = t * 32;
    const int rows = (n - j0 >= 32) ? 32 : (int)(n - j0);
    f32x4 v[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) v[u] = vn[u];
    const int64_t tn = t + (int64_t)B.gx * 4;
    load_tile((tn < tiles ? tn : t) * 32, vn);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int e = lane + 64 * u;
      if (e < npc) {
        const int rr = e / kq, q = e - rr * kq;
        const bool ok = rr < rows;
        float* d = zt + rr * P + 4 * q;
        d[0] = ok ? v[u][0] : 0.0f; d[1] = ok ? v[u][1] : 0.0f; d[2] = ok ? v[u][2] : 0.0f; d[3] = ok ? v[u][3] : 0.0f;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    f32x16 acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = zero16();
#pragma unroll
    for (int st = 0; st < KMAX / 2; ++st) {
      if (2 * st >= k4) break;
      const float a = zt[(lane & 31) * P + 2 * st + (lane >> 5)];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[nt] = mfma_f32(a, ub[nt][st], acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int v2 = 0; v2 < 16; ++v2) {
        const int rr = mfma32_row(v2, lane), c = 32 * nt + (lane & 31);
        if (rr < rows && c < r) out[(j0 + rr) * r + c] = acc[nt][v2];
      }
  }
}

// small n: RM lanes per output row (lane c computes column c), so short sweeps still fill the
// machine; the RM lanes of one row read the same Z row with 16-byte loads
template <int RM>
__global__ __launch_bounds__(256) void k_next_from_z(const float* __restrict__ Z, int64_t ldz, int64_t n, int k,
                                                     const float* __restrict__ Ur, int r, float* __restrict__ out) {
  next_rows_body<RM>(Z, ldz, n, k, Ur, r, out, hw_blk());
}

// us: (KMAX + 4) RM LDS floats
template <int RM>
__device__ __forceinline__ void next_cols_body(const float* __restrict__ Z, int64_t ldz, int64_t n, int k,
                                               const float* __restrict__ Ur, int r, float* __restrict__ out,
                                               float* us, const Blk B) {
  const int k4 = (k + 3) & ~3;
  for (int e = threadIdx.x; e < k4 * RM; e += 256) us[e] = Ur[e];   // Ur is k4 x RM, zero-padded
  __syncthreads();
  constexpr int ROWS = 256 / RM;
  const int jj = threadIdx.x / RM, c = threadIdx.x % RM;
  for (int64_t j = (int64_t)B.x * ROWS + jj; j < n; j += (int64_t)B.gx * ROWS) {
    const f32x4* z = reinterpret_cast<const f32x4*>(Z + j * ldz);
    float a0 = 0.0f, a1 = 0.0f;
    for (int q = 0; q < k4 / 4; ++q) {
      const f32x4 v = z[q];
      a0 += us[(4 * q) * RM + c] * v[0];
      a1 += us[(4 * q + 1) * RM + c] * v[1];
      a0 += us[(4 * q + 2) * RM + c] * v[2];
      a1 += us[(4 * q + 3) * RM + c] * v[3];
    }
    if (c < r) out[j * r + c] = a0 + a1;
  }
}
template <int RM>
__global__ __launch_bounds__(256) void k_next_from_z_cols(const float* __restrict__ Z, int64_t ldz, int64_t n,
                                                          int k, const float* __restrict__ Ur, int r,
                                                          float* __restrict__ out) {
  __shared__ float us[(KMAX + 4) * RM];
  next_cols_body<RM>(Z, ldz, n, k, Ur, r, out, us, hw_blk());
}
// both remainder layouts and every register width in one launch (jobs pick theirs)
__global__ __launch_bounds__(256) void k_next_grp(const JobNext* __restrict__ jobs, const int2* __restrict__ map) {
  __shared__ float us[(KMAX + 4) * 64];
  Blk B;
  const JobNext& j = grp_job(jobs, map, B);
  if (j.cols) {
    switch (j.rm) {
      case 8: next_cols_body<8>(j.Z, j.ldz, j.n, j.k, j.Ur, j.r, j.out, us, B); break;
      case 16: next_cols_body<16>(j.Z, j.ldz, j.n, j.k, j.Ur, j.r, j.out, us, B); break;
      case 32: next_cols_body<32>(j.Z, j.ldz, j.n, j.k, j.Ur, j.r, j.out, us, B); break;
      default: next_cols_body<64>(j.Z, j.ldz, j.n, j.k, j.Ur, j.r, j.out, us, B); break;
    }
  } else {
    switch (j.rm) {
      case 8: next_rows_body<8>(j.Z, j.ldz, j.n, j.k, j.Ur, j.r, j.out, B); break;
      case 16: next_rows_body<16>(j.Z, j.ldz, j.n, j.k, j.Ur, j.r, j.out, B); break;
      case 32: next_rows_body<32>(j.Z, j.ldz, j.n, j.k, j.Ur, j.r, j.out, B); break;
      default: next_rows_body<64>(j.Z, j.ldz, j.n, j.k, j.Ur, j.r, j.out, B); break;
    }
  }
}

int64_t next_from_z_grid(int64_t n, int r, bool* cols, bool batch) {
  const int RM = ur_ld(r);
  const int64_t cap = (int64_t)num_cus() * 16;
  // short sweeps: a row per lane would leave most of the machine idle, so RM lanes per row (not in a
  // batch, whose tensors together fill it)
  *cols = !batch && n < ((int64_t)1 << 17);
  int64_t grid = *cols ? (n * RM + 255) / 256 : (n + 255) / 256;
  if (grid > cap) grid = cap;
  if (grid < 1) grid = 1;
  return grid;
}

void launch_next_from_z(const float* Z, int64_t ldz, int64_t n, int k, const float* Ur, int r, float* out,
                        hipStream_t st, bool batch) {
  const int RM = ur_ld(r);
  bool cols = false;
  const int64_t grid = next_from_z_grid(n, r, &cols, batch);
  if (cols) {
#define TTX_NFZ(R) hipLaunchKernelGGL((k_next_from_z_cols<R>), dim3((unsigned)grid), dim3(256), 0, st, Z, ldz, n, k, Ur, r, out)
    if (RM == 8) TTX_NFZ(8); else if (RM == 16) TTX_NFZ(16); else if (RM == 32) TTX_NFZ(32); else TTX_NFZ(64);
#undef TTX_NFZ
    return;
  }
#define TTX_NFZ(R) hipLaunchKernelGGL((k_next_from_z<R>), dim3((unsigned)grid), dim3(256), 0, st, Z, ldz, n, k, Ur, r, out)
  if (RM == 8) TTX_NFZ(8); else if (RM == 16) TTX_NFZ(16); else if (RM == 32) TTX_NFZ(32); else TTX_NFZ(64);
#undef TTX_NFZ
}

// ------------------------------------------------------------------------------------------------
// pending full-rank left factor of a fused step: A = (I_N (x) G^T) S with G the previous step's
// square core (ms x ms, column-major).  Row (b, i) of A is sum_a G[a][b] row (a, i) of S.
//   mode 0: Y_A = P Y_S, Y row-major m x kp           (sketch results)
//   mode 1: Q_S = P^T Q_A, Q column-major kp x m       (projection operands)
//   mode 2: Q_A = P Q_S, Q column-major kp x m         (a basis taken on S, before the core assembly)
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_blockmix(const float* __restrict__ G, int ms, int m, int kp, int mode,
                                                  const float* __restrict__ X, float* __restrict__ Y) {
  __shared__ float gs[KMAX * KMAX];
  for (int e = threadIdx.x; e < ms * ms; e += 256) gs[e] = G[e];
  __syncthreads();
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < (int64_t)m * kp; e += (int64_t)gridDim.x * 256) {
    if (mode == 0) {
      const int row = (int)(e / kp), c = (int)(e % kp);
      const int b = row % ms, i0 = row - b;
      float acc = 0.0f;
      for (int a2 = 0; a2 < ms; ++a2) acc += gs[a2 + ms * b] * X[(int64_t)(i0 + a2) * kp + c];
      Y[e] = acc;
    } else if (mode == 1) {
      const int c = (int)(e / m), row = (int)(e % m);
      const int a2 = row % ms, i0 = row - a2;
      float acc = 0.0f;
      for (int b = 0; b < ms; ++b) acc += gs[a2 + ms * b] * X[(int64_t)c * m + i0 + b];
      Y[e] = acc;
    } else {
      const int c = (int)(e / m), row = (int)(e % m);
      const int b = row % ms, i0 = row - b;
      float acc = 0.0f;
      for (int a2 = 0; a2 < ms; ++a2) acc += gs[a2 + ms * b] * X[(int64_t)c * m + i0 + a2];
      Y[e] = acc;
    }
  }
}

void launch_blockmix(const float* G, int ms, int m, int kp, int mode, const float* X, float* Y, hipStream_t st) {
  int64_t grid = ((int64_t)m * kp + 255) / 256;
  if (grid > 2048) grid = 2048;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_blockmix, dim3((unsigned)grid), dim3(256), 0, st, G, ms, m, kp, mode, X, Y);
}

// C_A = P C_S P^T for the exact path of a fused step (m x m, FP64, m <= 64)
__global__ __launch_bounds__(256) void k_gram_mix(const float* __restrict__ G, int ms, int m,
                                                  const double* __restrict__ Cs, double* __restrict__ Ca) {
  __shared__ double gs[KMAX * KMAX];
  __shared__ double t[KMAX][KMAX + 1];
  for (int e = threadIdx.x; e < ms * ms; e += 256) gs[e] = (double)G[e];
  __syncthreads();
  // t = P C_S
  for (int e = threadIdx.x; e < m * m; e += 256) {
    const int row = e / m, col = e % m;
    const int b = row % ms, i0 = row - b;
    double acc = 0.0;
    for (int a2 = 0; a2 < ms; ++a2) acc += gs[a2 + ms * b] * Cs[(i0 + a2) * m + col];
    t[row][col] = acc;
  }
  __syncthreads();
  // C_A = t P^T
  for (int e = threadIdx.x; e < m * m; e += 256) {
    const int row = e / m, col = e % m;
    const int b = col % ms, j0 = col - b;
    double acc = 0.0;
    for (int a2 = 0; a2 < ms; ++a2) acc += t[row][j0 + a2] * gs[a2 + ms * b];
    Ca[row * m + col] = acc;
  }
}

void launch_gram_mix(const float* G, int ms, int m, const double* Cs, double* Ca, hipStream_t st) {
  hipLaunchKernelGGL(k_gram_mix, dim3(1), dim3(256), 0, st, G, ms, m, Cs, Ca);
}

// ------------------------------------------------------------------------------------------------
// a-posteriori range check with held-out Gaussian test columns Yt = A Omega_t (m x p, row-major):
// ||(I - Q Q^T) Yt||_F^2 and ||Yt||_F^2, both per 64-row block (Halko, Martinsson, Tropp 2011, 4.3).
// The residual is formed explicitly (Yt - Q C) so that it is not lost to cancellation.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_test_proj(const float* __restrict__ Qc, int64_t ldq, int m, int k,
                                                   const float* __restrict__ Yt, int p, double* __restrict__ part) {
  // the block's rows of Yt (RB64 x p, p <= 16) go to LDS once; thread (c, t) then sums its 64 rows in
  // four independent chains (the Q column segment is 64 consecutive floats)
  __shared__ double ys[RB64][17];
  const int64_t i0 = (int64_t)blockIdx.x * RB64;
  for (int e = threadIdx.x; e < RB64 * p; e += 256) {
    const int rr = e / p, t = e % p;
    ys[rr][t] = (i0 + rr < m) ? (double)Yt[(i0 + rr) * p + t] : 0.0;
  }
  __syncthreads();
  const int nr = (m - i0) < RB64 ? (int)(m - i0) : RB64;
  for (int e = threadIdx.x; e < k * p; e += 256) {
    const int c = e / p, t = e % p;
    const float* q = Qc + (int64_t)c * ldq + i0;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int rr = 0;
    for (; rr + 3 < nr; rr += 4) {
      s0 += (double)q[rr] * ys[rr][t];
      s1 += (double)q[rr + 1] * ys[rr + 1][t];
      s2 += (double)q[rr + 2] * ys[rr + 2][t];
      s3 += (double)q[rr + 3] * ys[rr + 3][t];
    }
    for (; rr < nr; ++rr) s0 += (double)q[rr] * ys[rr][t];
    part[(int64_t)blockIdx.x * k * p + e] = (s0 + s1) + (s2 + s3);
  }
}

__global__ __launch_bounds__(256) void k_test_resid(const float* __restrict__ Qc, int64_t ldq, int m, int k,
                                                    const float* __restrict__ Yt, int p,
                                                    const double* __restrict__ part, int64_t S,
                                                    double* __restrict__ out) {
  __shared__ double C[KMAX * 16];
  __shared__ double red[2][256];
  for (int e = threadIdx.x; e < k * p; e += 256) {
    double s = 0.0;
#pragma unroll 16
    for (int64_t q = 0; q < S; ++q) s += part[q * k * p + e];   // loads in flight together, added in order
    C[e] = s;
  }
  __syncthreads();
  const int64_t i0 = (int64_t)blockIdx.x * RB64;
  double r2 = 0.0, y2 = 0.0;
  for (int e = threadIdx.x; e < RB64 * p; e += 256) {
    const int rr = e / p, t = e % p;
    const int64_t i = i0 + rr;
    if (i >= m) continue;
    const double y = (double)Yt[i * p + t];
    // four independent chains over the basis columns
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
    int c = 0;
    for (; c + 3 < k; c += 4) {
      p0 += (double)Qc[(int64_t)c * ldq + i] * C[c * p + t];
      p1 += (double)Qc[(int64_t)(c + 1) * ldq + i] * C[(c + 1) * p + t];
      p2 += (double)Qc[(int64_t)(c + 2) * ldq + i] * C[(c + 2) * p + t];
      p3 += (double)Qc[(int64_t)(c + 3) * ldq + i] * C[(c + 3) * p + t];
    }
    for (; c < k; ++c) p0 += (double)Qc[(int64_t)c * ldq + i] * C[c * p + t];
    const double d = y - ((p0 + p1) + (p2 + p3));
    r2 += d * d;
    y2 += y * y;
  }
  red[0][threadIdx.x] = r2;
  red[1][threadIdx.x] = y2;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (threadIdx.x < s2) { red[0][threadIdx.x] += red[0][threadIdx.x + s2]; red[1][threadIdx.x] += red[1][threadIdx.x + s2]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[2 * blockIdx.x] = red[0][0]; out[2 * blockIdx.x + 1] = red[1][0]; }
}

void launch_test_resid(const float* Qc, int64_t ldq, int m, int k, const float* Yt, int p, double* part, double* out,
                       hipStream_t st) {
  const int64_t nb = (m + RB64 - 1) / RB64;
  hipLaunchKernelGGL(k_test_proj, dim3((unsigned)nb), dim3(256), 0, st, Qc, ldq, m, k, Yt, p, part);
  hipLaunchKernelGGL(k_test_resid, dim3((unsigned)nb), dim3(256), 0, st, Qc, ldq, m, k, Yt, p, part, nb, out);
}

// Yt (m x p row-major) <- columns c0 .. c0+p-1 of Y (m x kpy row-major)
__global__ void k_copy_block(const float* __restrict__ Y, int kpy, int m, int c0, int p, float* __restrict__ Yt) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)m * p; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / p;
    const int t = (int)(e % p);
    Yt[e] = Y[i * kpy + c0 + t];
  }
}

void launch_copy_block(const float* Y, int kpy, int m, int c0, int p, float* Yt, hipStream_t st) {
  int64_t grid = ((int64_t)m * p + 255) / 256;
  if (grid > 1024) grid = 1024;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_copy_block, dim3((unsigned)grid), dim3(256), 0, st, Y, kpy, m, c0, p, Yt);
}

__global__ void k_expand_rows(const float* __restrict__ compact, int r1c, int r1, int64_t N, int r,
                              float* __restrict__ full) {
  const int64_t total = (int64_t)r1 * N * r;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e % r1);
    const int64_t rest = e / r1;   // i + N c
    full[e] = (b < r1c) ? compact[b + (int64_t)r1c * rest] : 0.0f;
  }
}

void launch_expand_rows(const float* compact, int r1c, int r1, int64_t N, int r, float* full, hipStream_t st) {
  const int64_t total = (int64_t)r1 * N * r;
  int64_t grid = (total + 255) / 256;
  if (grid > (int64_t)num_cus() * 8) grid = (int64_t)num_cus() * 8;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_expand_rows, dim3((unsigned)grid), dim3(256), 0, st, compact, r1c, r1, N, r, full);
}

__device__ __forceinline__ void copy_cols_body(const float* __restrict__ A, int64_t lda, int m, int n,
                                               float* __restrict__ Y, int kp, const Blk B) {
  for (int64_t e = B.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)m * kp; e += (int64_t)B.gx * blockDim.x) {
    const int i = (int)(e / kp), c = (int)(e % kp);
    Y[e] = (c < n) ? A[i + (int64_t)c * lda] : 0.0f;
  }
}
__global__ void k_copy_cols(const float* __restrict__ A, int64_t lda, int m, int n, float* __restrict__ Y, int kp) {
  copy_cols_body(A, lda, m, n, Y, kp, hw_blk());
}
__global__ __launch_bounds__(256) void k_copy_cols_grp(const JobCopyCols* __restrict__ jobs, const int2* __restrict__ map) {
  Blk B;
  const JobCopyCols& j = grp_job(jobs, map, B);
  copy_cols_body(j.A, j.lda, j.m, j.n, j.Y, j.kp, B);
}

int64_t copy_cols_grid(int m, int kp) {
  int64_t grid = ((int64_t)m * kp + 255) / 256;
  if (grid > 4096) grid = 4096;
  if (grid < 1) grid = 1;
  return grid;
}

void launch_copy_cols(const float* A, int64_t lda, int m, int n, float* Y, int kp, hipStream_t st) {
  const int64_t grid = copy_cols_grid(m, kp);
  hipLaunchKernelGGL(k_copy_cols, dim3((unsigned)grid), dim3(256), 0, st, A, lda, m, n, Y, kp);
}

// ------------------------------------------------------------------------------------------------
// grouped launchers
// ------------------------------------------------------------------------------------------------
#define TTX_GRP_ARGS dim3((unsigned)g.blocks), dim3(g.threads), 0, st
int grp_kcap(int k) { return k <= 32 ? 32 : (k <= 48 ? 48 : 64); }
void grp_gram_rows(bool f64, int kc, const GroupLaunch& g, hipStream_t st) {
  const JobGramRows* J = static_cast<const JobGramRows*>(g.jobs);
#define TTX_G(KC) { if (f64) hipLaunchKernelGGL((k_gram_rows_grp<double, KC>), TTX_GRP_ARGS, J, g.map); \
                    else hipLaunchKernelGGL((k_gram_rows_grp<float, KC>), TTX_GRP_ARGS, J, g.map); }
  if (kc <= 32) TTX_G(32) else if (kc <= 48) TTX_G(48) else TTX_G(64)
#undef TTX_G
}
void grp_chol(const GroupLaunch& g, hipStream_t st) {
  hipLaunchKernelGGL(k_chol_grp, TTX_GRP_ARGS, static_cast<const JobChol*>(g.jobs), g.map);
}
void grp_apply(int kind, int kc, const GroupLaunch& g, hipStream_t st) {
  const JobApply* J = static_cast<const JobApply*>(g.jobs);
#define TTX_G(KC) { if (kind == 0) hipLaunchKernelGGL((k_apply_grp<float, 0, KC>), TTX_GRP_ARGS, J, g.map); \
                    else if (kind == 1) hipLaunchKernelGGL((k_apply_grp<double, 0, KC>), TTX_GRP_ARGS, J, g.map); \
                    else hipLaunchKernelGGL((k_apply_grp<double, 1, KC>), TTX_GRP_ARGS, J, g.map); }
  if (kc <= 32) TTX_G(32) else if (kc <= 48) TTX_G(48) else TTX_G(64)
#undef TTX_G
}
void grp_syevj(const GroupLaunch& g, hipStream_t st, int max_sweeps) {
  hipLaunchKernelGGL(k_syevj_grp, TTX_GRP_ARGS, static_cast<const JobSyevj*>(g.jobs), g.map, max_sweeps);
}
void grp_assemble(const GroupLaunch& g, hipStream_t st) {
  const JobAssemble* J = static_cast<const JobAssemble*>(g.jobs);
  hipLaunchKernelGGL(k_assemble_g_grp, TTX_GRP_ARGS, J, g.map);
  hipLaunchKernelGGL(k_assemble_fix_grp, TTX_GRP_ARGS, J, g.map);
}
void grp_next(const GroupLaunch& g, hipStream_t st) {
  hipLaunchKernelGGL(k_next_grp, TTX_GRP_ARGS, static_cast<const JobNext*>(g.jobs), g.map);
}
void grp_copy_cols(const GroupLaunch& g, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_cols_grp, TTX_GRP_ARGS, static_cast<const JobCopyCols*>(g.jobs), g.map);
}
#undef TTX_GRP_ARGS

}  // namespace ttx
