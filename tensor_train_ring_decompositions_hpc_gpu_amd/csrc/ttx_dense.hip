// FP64 small dense kernels of one decomposition step (k = sketch width <= 64):
//   Gram of tall panels, shifted Cholesky + explicit triangular inverse (CholeskyQR rounds),
//   cyclic Jacobi eigensolver for the projected Gram, core assembly with the sign gauge, and the
//   remainder A_next = U_r^T Z^T.
// These replace the reference's cuSOLVER calls: basicQR (geqrf + orgqr,
// ttdecomposition/third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:85-159), basicEig
// (syevd, ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:334-385) and the descending
// column selection matrixInvertColumn (:274-290).
#include "ttx_device.h"
#include "ttx_kernels.h"

namespace ttx {

constexpr int KMAX = 64;

// ------------------------------------------------------------------------------------------------
// Gram of a tall panel: part[blockIdx.x][a][b] = sum over the WG's rows of X[i][a] X[i][b]
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_gram_tall(const T* __restrict__ X, int64_t ldx, int64_t rows, int k,
                                                   int64_t rows_per_wg, double* __restrict__ part) {
  __shared__ double xs[32][KMAX + 1];
  const int kb = (k + 3) / 4;
  const int nb = kb * (kb + 1) / 2;
  const int tid = threadIdx.x;
  // up to 2 blocks of 4x4 per thread (nb <= 136 for k <= 64)
  int ta[2], tb[2];
  int nmine = 0;
  for (int bi = tid, q = 0; bi < nb && q < 2; bi += 256, ++q) {
    // bi -> (a, b) with a <= b, row-major over the upper triangle
    int a = 0, rem = bi;
    while (rem >= kb - a) { rem -= kb - a; ++a; }
    ta[q] = a;
    tb[q] = a + rem;
    nmine = q + 1;
  }
  double acc[2][16];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[q][e] = 0.0;

  const int64_t r0 = (int64_t)blockIdx.x * rows_per_wg;
  const int64_t r1 = (r0 + rows_per_wg < rows) ? r0 + rows_per_wg : rows;
  for (int64_t rc = r0; rc < r1; rc += 32) {
    for (int e = tid; e < 32 * kb * 4; e += 256) {
      const int rr = e / (kb * 4), c = e % (kb * 4);
      const int64_t i = rc + rr;
      xs[rr][c] = (i < r1 && c < k) ? (double)X[i * ldx + c] : 0.0;
    }
    __syncthreads();
    for (int q = 0; q < nmine; ++q) {
      const int ca = 4 * ta[q], cb = 4 * tb[q];
      for (int rr = 0; rr < 32; ++rr) {
        double xa[4], xb[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { xa[u] = xs[rr][ca + u]; xb[u] = xs[rr][cb + u]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[q][4 * u + v] += xa[u] * xb[v];
      }
    }
    __syncthreads();
  }
  double* o = part + (int64_t)blockIdx.x * k * k;
  for (int q = 0; q < nmine; ++q) {
    for (int u = 0; u < 4; ++u)
      for (int v = 0; v < 4; ++v) {
        const int a = 4 * ta[q] + u, b = 4 * tb[q] + v;
        if (a < k && b < k) {
          o[a * k + b] = acc[q][4 * u + v];
          o[b * k + a] = acc[q][4 * u + v];
        }
      }
  }
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
  rpw = ((rpw + 31) / 32) * 32;
  hipLaunchKernelGGL(k_gram_tall<float>, dim3((unsigned)grid), dim3(256), 0, st, X, ldx, rows, k, rpw, part);
}
void launch_gram_tall_f64(const double* X, int64_t ldx, int64_t rows, int k, double* part, int64_t grid,
                          hipStream_t st) {
  int64_t rpw = (rows + grid - 1) / grid;
  rpw = ((rpw + 31) / 32) * 32;
  hipLaunchKernelGGL(k_gram_tall<double>, dim3((unsigned)grid), dim3(256), 0, st, X, ldx, rows, k, rpw, part);
}

// ------------------------------------------------------------------------------------------------
// Cholesky (+ optional shift) and explicit inverse of the upper factor
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_chol_inv(const double* __restrict__ G, int k, int64_t m, int shift_mode,
                                                  double* __restrict__ Rinv, int* __restrict__ status) {
  __shared__ double a[KMAX][KMAX + 1];
  __shared__ double R[KMAX][KMAX + 1];
  __shared__ double shift_s;
  __shared__ double maxdiag;
  __shared__ int nzero;
  const int tid = threadIdx.x;
  for (int e = tid; e < k * k; e += 256) {
    const int i = e / k, j = e % k;
    a[i][j] = G[e];
    R[i][j] = 0.0;
  }
  __syncthreads();
  if (tid == 0) {
    double tr = 0.0, md = 0.0;
    for (int i = 0; i < k; ++i) { tr += a[i][i]; md = fmax(md, a[i][i]); }
    // shifted CholeskyQR (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa 2020):
    // s = 11 (m k + k (k+1)) u ||X||_2^2 with ||X||_2^2 <= trace(X^T X)
    shift_s = shift_mode ? 11.0 * ((double)m * k + (double)k * (k + 1)) * 1.1102230246251565e-16 * tr : 0.0;
    maxdiag = md + shift_s;
    nzero = 0;
  }
  __syncthreads();
  if (shift_mode)
    for (int i = tid; i < k; i += 256) a[i][i] += shift_s;
  __syncthreads();
  const double tiny = maxdiag * 1e-28;
  for (int j = 0; j < k; ++j) {
    const double d = a[j][j];
    const bool ok = d > tiny;
    const double rjj = ok ? sqrt(d) : 0.0;
    for (int c = j + tid; c < k; c += 256) R[j][c] = ok ? (c == j ? rjj : a[j][c] / rjj) : 0.0;
    if (!ok && tid == 0) nzero += 1;
    __syncthreads();
    // trailing update a[p][q] -= R[j][p] R[j][q], j < p <= q
    const int t = k - j - 1;
    for (int e = tid; e < t * t; e += 256) {
      const int p = j + 1 + e / t, q = j + 1 + e % t;
      if (q >= p) a[p][q] -= R[j][p] * R[j][q];
    }
    __syncthreads();
  }
  // Rinv: R Rinv = I, one column per thread (column c of `a` holds x); zero pivots give zero rows /
  // columns
  for (int c = tid; c < k; c += 256) {
    for (int i = 0; i < k; ++i) a[i][c] = 0.0;
    if (R[c][c] != 0.0) {
      a[c][c] = 1.0 / R[c][c];
      for (int i = c - 1; i >= 0; --i) {
        if (R[i][i] == 0.0) continue;
        double s = 0.0;
        for (int l = i + 1; l <= c; ++l) s += R[i][l] * a[l][c];
        a[i][c] = -s / R[i][i];
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < k * k; e += 256) Rinv[e] = a[e / k][e % k];
  if (tid == 0 && status) status[0] += nzero;
}

void launch_chol_inv(const double* G, int k, int64_t m, int shift_mode, double* Rinv, int* status, hipStream_t st) {
  hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(256), 0, st, G, k, m, shift_mode, Rinv, status);
}

// ------------------------------------------------------------------------------------------------
// W_out = W_in Rinv (rows x k), Rinv upper triangular (row-major k x k)
// ------------------------------------------------------------------------------------------------
template <typename TI, int OUT>   // OUT 0: double row-major (stride k); OUT 1: float column-major Qc
__global__ __launch_bounds__(256) void k_apply_rinv(const TI* __restrict__ Win, int64_t ldi, int64_t rows, int k,
                                                    int kp, const double* __restrict__ Rinv, double* __restrict__ Wd,
                                                    float* __restrict__ Qc, int64_t ldq) {
  __shared__ double ri[KMAX][KMAX + 1];
  __shared__ double ws[64][KMAX + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < k * k; e += 256) ri[e / k][e % k] = Rinv[e];
  const int64_t i0 = (int64_t)blockIdx.x * 64;
  for (int e = tid; e < 64 * k; e += 256) {
    const int rr = e / k, c = e % k;
    const int64_t i = i0 + rr;
    ws[rr][c] = (i < rows) ? (double)Win[i * ldi + c] : 0.0;
  }
  __syncthreads();
  const int KPo = OUT ? kp : k;
  for (int e = tid; e < 64 * KPo; e += 256) {
    // OUT 1: consecutive threads take consecutive rows of one column (coalesced column-major store)
    const int rr = OUT ? (e % 64) : (e / KPo);
    const int c = OUT ? (e / 64) : (e % KPo);
    const int64_t i = i0 + rr;
    if (i >= rows) continue;
    double s = 0.0;
    if (c < k)
      for (int a2 = 0; a2 <= c; ++a2) s += ws[rr][a2] * ri[a2][c];
    if (OUT) Qc[(int64_t)c * ldq + i] = (float)s;
    else Wd[i * k + c] = s;
  }
}

void launch_apply_rinv_f32(const float* Win, int64_t ldi, int64_t rows, int k, const double* Rinv, double* Wout,
                           hipStream_t st) {
  const int64_t grid = (rows + 63) / 64;
  hipLaunchKernelGGL((k_apply_rinv<float, 0>), dim3((unsigned)grid), dim3(256), 0, st, Win, ldi, rows, k, k, Rinv,
                     Wout, (float*)nullptr, (int64_t)0);
}
void launch_apply_rinv_f64(const double* Win, int64_t rows, int k, const double* Rinv, double* Wout, hipStream_t st) {
  const int64_t grid = (rows + 63) / 64;
  hipLaunchKernelGGL((k_apply_rinv<double, 0>), dim3((unsigned)grid), dim3(256), 0, st, Win, (int64_t)k, rows, k, k,
                     Rinv, Wout, (float*)nullptr, (int64_t)0);
}
void launch_apply_rinv_out(const double* Win, int64_t rows, int k, int kp, const double* Rinv, float* Qc,
                           int64_t ldq, hipStream_t st) {
  const int64_t grid = (rows + 63) / 64;
  hipLaunchKernelGGL((k_apply_rinv<double, 1>), dim3((unsigned)grid), dim3(256), 0, st, Win, (int64_t)k, rows, k, kp,
                     Rinv, (double*)nullptr, Qc, ldq);
}

// ------------------------------------------------------------------------------------------------
// cyclic Jacobi eigensolver (round-robin ordering), FP64, k <= 64
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_syevj(const double* __restrict__ G, int k, double* __restrict__ lam,
                                               double* __restrict__ Vout, int* __restrict__ status) {
  __shared__ double a[KMAX][KMAX + 1];
  __shared__ double v[KMAX][KMAX + 1];
  __shared__ double cs[KMAX / 2], sn[KMAX / 2];
  __shared__ int pp[KMAX / 2], qq[KMAX / 2];
  __shared__ double red[256];
  __shared__ int done;
  const int tid = threadIdx.x;
  const int k2 = k + (k & 1);   // an odd size gets one zero row / column (its pairs rotate by identity)
  for (int e = tid; e < k2 * k2; e += 256) {
    const int i = e / k2, j = e % k2;
    a[i][j] = (i < k && j < k) ? G[i * k + j] : 0.0;
    v[i][j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int np = k2 / 2;
  int sweeps = 0;
  for (; sweeps < 40; ++sweeps) {
    // convergence: off-diagonal mass relative to the total
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < k2 * k2; e += 256) {
      const int i = e / k2, j = e % k2;
      const double x = a[i][j] * a[i][j];
      tot += x;
      if (i != j) off += x;
    }
    red[tid] = off;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    const double offs = red[0];
    __syncthreads();
    red[tid] = tot;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    if (tid == 0) done = (offs <= 1e-30 * red[0]) || red[0] == 0.0;
    __syncthreads();
    if (done) break;
    for (int rnd = 0; rnd < k2 - 1; ++rnd) {
      if (tid < np) {
        int x, y;
        if (tid == 0) { x = rnd % (k2 - 1); y = k2 - 1; }
        else { x = (rnd + tid) % (k2 - 1); y = (rnd - tid + (k2 - 1)) % (k2 - 1); }
        const int p = x < y ? x : y, q = x < y ? y : x;
        const double apq = a[p][q];
        double c = 1.0, s = 0.0;
        if (apq != 0.0) {
          const double tau = (a[q][q] - a[p][p]) / (2.0 * apq);
          const double t = (tau >= 0.0) ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          s = t * c;
        }
        pp[tid] = p; qq[tid] = q; cs[tid] = c; sn[tid] = s;
      }
      __syncthreads();
      // rows: J^T A
      for (int e = tid; e < np * k2; e += 256) {
        const int i = e / k2, j = e % k2;
        const int p = pp[i], q = qq[i];
        const double c = cs[i], s = sn[i];
        const double ap = a[p][j], aq = a[q][j];
        a[p][j] = c * ap - s * aq;
        a[q][j] = s * ap + c * aq;
      }
      __syncthreads();
      // columns: (J^T A) J and V J
      for (int e = tid; e < np * k2; e += 256) {
        const int i = e / k2, j = e % k2;
        const int p = pp[i], q = qq[i];
        const double c = cs[i], s = sn[i];
        const double ap = a[j][p], aq = a[j][q];
        a[j][p] = c * ap - s * aq;
        a[j][q] = s * ap + c * aq;
        const double vp = v[j][p], vq = v[j][q];
        v[j][p] = c * vp - s * vq;
        v[j][q] = s * vp + c * vq;
      }
      __syncthreads();
    }
  }
  // sort descending: rank of each eigenvalue (ties broken by index)
  for (int i = tid; i < k; i += 256) {
    const double li = a[i][i];
    int rank = 0;
    for (int j = 0; j < k; ++j) {
      const double lj = a[j][j];
      if (lj > li || (lj == li && j < i)) ++rank;
    }
    lam[rank] = li;
    for (int r = 0; r < k; ++r) Vout[r * k + rank] = v[r][i];
  }
  if (tid == 0 && status) status[1] = sweeps;
}

void launch_syevj(const double* G, int k, double* lam, double* V, int* status, double* work, hipStream_t st) {
  (void)work;
  hipLaunchKernelGGL(k_syevj, dim3(1), dim3(256), 0, st, G, k, lam, V, status);
}

// ------------------------------------------------------------------------------------------------
// core assembly with the sign gauge
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_assemble(const float* __restrict__ Qc, int64_t ldq, int m, int k,
                                                   const double* __restrict__ lam, const double* __restrict__ V,
                                                   int r, float* __restrict__ Gc, float* __restrict__ Ur,
                                                   float* __restrict__ Qn, int kpn, float* __restrict__ sigma) {
  __shared__ double vs[KMAX][KMAX + 1];
  __shared__ float bestv[1024];
  __shared__ int besti[1024];
  __shared__ float sgn[128];
  const int tid = threadIdx.x;
  const int rk = r < k ? r : k;   // columns beyond k are zero
  for (int e = tid; e < k * k; e += 1024) vs[e / k][e % k] = V[e];
  __syncthreads();
  // pass 1: G = Q V[:, :rk] into Gc (column-major m x r), columns rk..r-1 zero
  for (int64_t e = tid; e < (int64_t)m * r; e += 1024) {
    const int c = (int)(e / m), i = (int)(e % m);
    double s = 0.0;
    if (c < rk) {
      if (Qc) {
        for (int a2 = 0; a2 < k; ++a2) s += (double)Qc[(int64_t)a2 * ldq + i] * vs[a2][c];
      } else {
        s = vs[i][c];
      }
    }
    Gc[(int64_t)c * m + i] = (float)s;
  }
  __syncthreads();
  // pass 2: per column, the first index of the largest |entry| decides the sign
  for (int c = 0; c < r; ++c) {
    float bv = -1.0f;
    int bi = 0x7fffffff;
    for (int i = tid; i < m; i += 1024) {
      const float x = fabsf(Gc[(int64_t)c * m + i]);
      if (x > bv) { bv = x; bi = i; }
    }
    bestv[tid] = bv;
    besti[tid] = bi;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
      if (tid < s) {
        const float ov = bestv[tid + s];
        const int oi = besti[tid + s];
        if (ov > bestv[tid] || (ov == bestv[tid] && oi < besti[tid])) { bestv[tid] = ov; besti[tid] = oi; }
      }
      __syncthreads();
    }
    if (tid == 0) {
      const int i = besti[0];
      sgn[c] = (i < m && Gc[(int64_t)c * m + i] < 0.0f) ? -1.0f : 1.0f;
    }
    __syncthreads();
  }
  // pass 3: apply signs, write Ur, Qn, sigma
  for (int64_t e = tid; e < (int64_t)m * r; e += 1024) {
    const int c = (int)(e / m);
    Gc[e] *= sgn[c];
  }
  __syncthreads();
  if (Ur)
    for (int e = tid; e < k * r; e += 1024) {
      const int a2 = e / r, c = e % r;
      Ur[e] = (c < rk) ? (float)(vs[a2][c] * (double)sgn[c]) : 0.0f;
    }
  if (Qn)
    for (int64_t e = tid; e < (int64_t)m * kpn; e += 1024) {
      const int c = (int)(e / m), i = (int)(e % m);
      Qn[e] = (c < r) ? Gc[(int64_t)c * m + i] : 0.0f;
    }
  if (sigma)
    for (int i = tid; i < k; i += 1024) sigma[i] = (float)sqrt(fmax(lam[i], 0.0));
}

void launch_assemble(const float* Qc, int64_t ldq, int m, int k, const double* lam, const double* V, int r, float* Gc,
                     float* Ur, float* Qn, int kpn, float* sigma, hipStream_t st) {
  hipLaunchKernelGGL(k_assemble, dim3(1), dim3(1024), 0, st, Qc, ldq, m, k, lam, V, r, Gc, Ur, Qn, kpn, sigma);
}

// ------------------------------------------------------------------------------------------------
// A_next = Ur^T Z^T  (r x n column-major)
// ------------------------------------------------------------------------------------------------
template <int RM>
__global__ __launch_bounds__(256) void k_next_from_z(const float* __restrict__ Z, int64_t ldz, int64_t n, int k,
                                                     const float* __restrict__ Ur, int r, float* __restrict__ out) {
  __shared__ float us[KMAX * RM];
  for (int e = threadIdx.x; e < k * r; e += 256) us[(e / r) * RM + (e % r)] = Ur[e];
  for (int e = threadIdx.x; e < k * RM; e += 256)
    if ((e % RM) >= r) us[e] = 0.0f;
  __syncthreads();
  for (int64_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    float acc[RM];
#pragma unroll
    for (int c = 0; c < RM; ++c) acc[c] = 0.0f;
    const float* z = Z + j * ldz;
    for (int a2 = 0; a2 < k; ++a2) {
      const float zv = z[a2];
#pragma unroll
      for (int c = 0; c < RM; ++c) acc[c] += us[a2 * RM + c] * zv;
    }
    float* o = out + j * r;
#pragma unroll
    for (int c = 0; c < RM; ++c)
      if (c < r) o[c] = acc[c];
  }
}

void launch_next_from_z(const float* Z, int64_t ldz, int64_t n, int k, const float* Ur, int r, float* out,
                        hipStream_t st) {
  int64_t grid = (n + 255) / 256;
  const int64_t cap = (int64_t)num_cus() * 8;
  if (grid > cap) grid = cap;
  if (grid < 1) grid = 1;
  if (r <= 8) hipLaunchKernelGGL((k_next_from_z<8>), dim3((unsigned)grid), dim3(256), 0, st, Z, ldz, n, k, Ur, r, out);
  else if (r <= 16) hipLaunchKernelGGL((k_next_from_z<16>), dim3((unsigned)grid), dim3(256), 0, st, Z, ldz, n, k, Ur, r, out);
  else if (r <= 32) hipLaunchKernelGGL((k_next_from_z<32>), dim3((unsigned)grid), dim3(256), 0, st, Z, ldz, n, k, Ur, r, out);
  else hipLaunchKernelGGL((k_next_from_z<64>), dim3((unsigned)grid), dim3(256), 0, st, Z, ldz, n, k, Ur, r, out);
}

__global__ void k_copy_cols(const float* __restrict__ A, int64_t lda, int m, int n, float* __restrict__ Y, int kp) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)m * kp;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / kp), c = (int)(e % kp);
    Y[e] = (c < n) ? A[i + (int64_t)c * lda] : 0.0f;
  }
}

void launch_copy_cols(const float* A, int64_t lda, int m, int n, float* Y, int kp, hipStream_t st) {
  int64_t grid = ((int64_t)m * kp + 255) / 256;
  if (grid > 4096) grid = 4096;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_copy_cols, dim3((unsigned)grid), dim3(256), 0, st, A, lda, m, n, Y, kp);
}

}  // namespace ttx
