// Column-major FP32 GEMM for core contractions (synthetic inputs and reconstructions), and the
// fused residual form that never writes the reconstruction:
//   C = alpha A B + beta C,  or  per block: sum (X - A B)^2 and sum X^2 in FP64.
// This is the build's counterpart of the host-staged SGEMM chains smallargeMatrixMulti_slice_host
// used by generalTTTensor and calMSE (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:75-182);
// unlike calMSE it leaves the input tensor untouched.
#include "ttx_device.h"
#include "ttx_kernels.h"

namespace ttx {

constexpr int BM = 64, BN = 64, BK = 16;

template <bool RESID>
__global__ __launch_bounds__(256) void k_gemm(int64_t M, int64_t N, int64_t K, float alpha, const float* __restrict__ A,
                                              int64_t lda, const float* __restrict__ B, int64_t ldb, float beta,
                                              float* __restrict__ C, int64_t ldc, double* __restrict__ resid) {
  __shared__ float As[BK][BM + 4];
  __shared__ float Bs[BK][BN + 4];
  __shared__ double red[2][256];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  const int64_t nbm = (M + BM - 1) / BM;
  const int64_t blk = blockIdx.x;
  const int64_t m0 = (blk % nbm) * BM, n0 = (blk / nbm) * BN;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    // A tile BM x BK: thread loads 4 elements along M
    for (int e = tid; e < BM * BK; e += 256) {
      const int mm = e % BM, kk = e / BM;
      const int64_t gi = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gi < M && gk < K) ? A[gi + gk * lda] : 0.0f;
    }
    for (int e = tid; e < BN * BK; e += 256) {
      const int kk = e % BK, nn = e / BK;
      const int64_t gk = k0 + kk, gj = n0 + nn;
      Bs[kk][nn] = (gk < K && gj < N) ? B[gk + gj * ldb] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
  double r2 = 0.0, x2 = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t gi = m0 + tx + 16 * i, gj = n0 + ty + 16 * j;
      if (gi < M && gj < N) {
        float* cp = C + gi + gj * ldc;
        if (RESID) {
          const double x = (double)*cp;
          const double d = x - (double)acc[i][j];
          r2 += d * d;
          x2 += x * x;
        } else {
          const float v = alpha * acc[i][j];
          *cp = (beta == 0.0f) ? v : v + beta * *cp;
        }
      }
    }
  if (RESID) {
    red[0][tid] = r2;
    red[1][tid] = x2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) { red[0][tid] += red[0][tid + s]; red[1][tid] += red[1][tid + s]; }
      __syncthreads();
    }
    if (tid == 0) {
      resid[2 * blk] = red[0][0];
      resid[2 * blk + 1] = red[1][0];
    }
  }
}

int64_t gemm_resid_blocks(int64_t M, int64_t N) { return ((M + BM - 1) / BM) * ((N + BN - 1) / BN); }

void launch_gemm(int64_t M, int64_t N, int64_t K, float alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
                 float beta, float* C, int64_t ldc, hipStream_t st) {
  dim3 grid((unsigned)gemm_resid_blocks(M, N));
  hipLaunchKernelGGL((k_gemm<false>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                     (double*)nullptr);
}

void launch_gemm_resid(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda, const float* B, int64_t ldb,
                       const float* X, int64_t ldx, double* resid, hipStream_t st) {
  dim3 grid((unsigned)gemm_resid_blocks(M, N));
  hipLaunchKernelGGL((k_gemm<true>), grid, dim3(256), 0, st, M, N, K, 1.0f, A, lda, B, ldb, 0.0f,
                     const_cast<float*>(X), ldx, resid);
}

}  // namespace ttx
