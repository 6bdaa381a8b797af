// Seeded synthetic data: Philox normals, additive noise, sums of squares, and the explicit sketch
// matrix (the streaming kernels generate the same sketch entries in-kernel).
// The synthetic inputs replace the reference's glibc-rand generators generalTTTensor /
// generalTRTensor (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:185-271,
// trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:231-329) with the distributions of the
// benchmark workloads; element e of a filled array is normal4(seed, stream, e >> 2, 0)[e & 3].
#include "ttx_device.h"
#include "ttx_kernels.h"

namespace ttx {

__global__ void k_fill_normal(float* __restrict__ out, int64_t count, uint64_t seed, uint32_t stream, float scale) {
#pragma clang fp contract(off)
  const int64_t groups = (count + 3) / 4;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < groups; g += (int64_t)gridDim.x * blockDim.x) {
    float z[4];
    normal4(seed, stream, (uint64_t)g, 0u, z);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t i = 4 * g + e;
      if (i < count) out[i] = z[e] * scale;
    }
  }
}

__global__ void k_add_noise(float* __restrict__ A, int64_t count, uint64_t seed, uint32_t stream, float coef) {
#pragma clang fp contract(off)
  const int64_t groups = (count + 3) / 4;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < groups; g += (int64_t)gridDim.x * blockDim.x) {
    float z[4];
    normal4(seed, stream, (uint64_t)g, 0u, z);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t i = 4 * g + e;
      if (i < count) {
        const float t = coef * z[e];
        A[i] = A[i] + t;
      }
    }
  }
}

__global__ void k_sumsq(const float* __restrict__ A, int64_t count, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const double x = (double)A[i];
    s += x * x;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Omega (n x ell, row-major with stride ell): entry (j, c) = omega_round(normal4(seed, stream, j, c/4)[c%4])
__global__ void k_omega(float* __restrict__ out, int64_t n, int ell, uint64_t seed, uint32_t stream) {
  const int ng = (ell + 3) / 4;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * ng; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / ng;
    const int g = (int)(e % ng);
    float z[4];
    normal4(seed, stream, (uint64_t)j, (uint32_t)g, z);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int c = 4 * g + t;
      if (c < ell) out[j * ell + c] = omega_round(z[t]);
    }
  }
}

static unsigned grid_for(int64_t work, int64_t per_block, int64_t cap) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void launch_fill_normal(float* out, int64_t count, uint64_t seed, uint32_t stream, float scale, hipStream_t st) {
  const unsigned g = grid_for((count + 3) / 4, 256, (int64_t)num_cus() * 16);
  hipLaunchKernelGGL(k_fill_normal, dim3(g), dim3(256), 0, st, out, count, seed, stream, scale);
}

void launch_add_noise(float* A, int64_t count, uint64_t seed, uint32_t stream, float coef, hipStream_t st) {
  const unsigned g = grid_for((count + 3) / 4, 256, (int64_t)num_cus() * 16);
  hipLaunchKernelGGL(k_add_noise, dim3(g), dim3(256), 0, st, A, count, seed, stream, coef);
}

int64_t sumsq_grid(int64_t count) { return grid_for(count, 256 * 16, (int64_t)num_cus() * 8); }

void launch_sumsq(const float* A, int64_t count, double* out2, double* part, int64_t grid, hipStream_t st) {
  hipLaunchKernelGGL(k_sumsq, dim3((unsigned)grid), dim3(256), 0, st, A, count, part);
  launch_reduce_f64(part, grid, 1, out2, st);
}

void launch_omega(float* out, int64_t n, int ell, uint64_t seed, uint32_t stream, hipStream_t st) {
  const unsigned g = grid_for(n * ((ell + 3) / 4), 256, (int64_t)num_cus() * 16);
  hipLaunchKernelGGL(k_omega, dim3(g), dim3(256), 0, st, out, n, ell, seed, stream);
}

// zero-fill (a kernel, so that captured graphs hold no memset nodes)
__global__ void k_zero_words(uint32_t* __restrict__ p, int64_t words) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < words; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = 0u;
}

void launch_zero(void* p, int64_t bytes, hipStream_t st) {
  const int64_t words = bytes / 4;
  if (words <= 0) return;
  int64_t g = (words + 255) / 256;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)g), dim3(256), 0, st, reinterpret_cast<uint32_t*>(p), words);
}

}  // namespace ttx
