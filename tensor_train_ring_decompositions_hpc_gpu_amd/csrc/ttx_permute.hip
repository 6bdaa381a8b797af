// Bit-exact mode permutations of column-major tensors.
//
// permute_231 is the tensor-ring first-step relayout tensorToMode231
// (trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:49-58, launched at :873):
//   T2[b + r1 j + r1 n a] = T1[a + r0 b + r0 r1 j],  a < r0, b < r1, j < n,
// i.e. the transpose of T1 viewed as an r0 x (r1 n) column-major matrix.  Indices are 64-bit (the
// reference's int arithmetic overflows at r0 r1 n >= 2^31).  Each thread moves a 4 x 4 block
// (four consecutive columns p of T1, four values of a) with one 16-byte load per column and one
// 16-byte store per a, so both sides are fully coalesced.
#include "ttx_device.h"
#include "ttx_kernels.h"

namespace ttx {

__global__ __launch_bounds__(256) void k_permute_231_v4(const float* __restrict__ in, float* __restrict__ out,
                                                        int64_t r0, int64_t np) {
  // r0 % 4 == 0, np % 4 == 0; thread handles columns p .. p+3 and values a .. a+3
  const int64_t nq = np / 4;
  const int64_t na = r0 / 4;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nq * na; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = e % nq, ag = e / nq;
    const int64_t p = 4 * q, a = 4 * ag;
    f32x4 col[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) col[u] = *reinterpret_cast<const f32x4*>(in + (p + u) * r0 + a);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 v;
      v[0] = col[0][t]; v[1] = col[1][t]; v[2] = col[2][t]; v[3] = col[3][t];
      *reinterpret_cast<f32x4*>(out + (a + t) * np + p) = v;
    }
  }
}

__global__ void k_permute_231_s(const float* __restrict__ in, float* __restrict__ out, int64_t r0, int64_t np) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < r0 * np; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = e % np, a = e / np;   // consecutive threads write consecutive p
    out[e] = in[p * r0 + a];
  }
}

void launch_permute_231(const float* in, float* out, int64_t r0, int64_t r1, int64_t n, hipStream_t st) {
  const int64_t np = r1 * n;
  const bool aligned = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
  if (r0 % 4 == 0 && np % 4 == 0 && aligned) {
    int64_t g = (np / 4 * (r0 / 4) + 255) / 256;
    const int64_t cap = (int64_t)num_cus() * 32;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_permute_231_v4, dim3((unsigned)g), dim3(256), 0, st, in, out, r0, np);
  } else {
    int64_t g = (r0 * np + 255) / 256;
    const int64_t cap = (int64_t)num_cus() * 32;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_permute_231_s, dim3((unsigned)g), dim3(256), 0, st, in, out, r0, np);
  }
}

// generic: out has dims (dims[perm[0]], ..., dims[perm[nd-1]]); out(i_perm) = in(i)
struct PermArgs {
  int nd;
  int64_t idims[6];
  int64_t ostride[6];   // output stride of input mode d
};

__global__ void k_permute(const float* __restrict__ in, float* __restrict__ out, int64_t total, PermArgs pa) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t rem = e, o = 0;
    for (int d = 0; d < pa.nd; ++d) {
      const int64_t i = rem % pa.idims[d];
      rem /= pa.idims[d];
      o += i * pa.ostride[d];
    }
    out[o] = in[e];
  }
}

void launch_permute(const float* in, float* out, int nd, const int64_t* dims, const int* perm, hipStream_t st) {
  PermArgs pa;
  pa.nd = nd;
  int64_t total = 1;
  for (int d = 0; d < nd; ++d) { pa.idims[d] = dims[d]; total *= dims[d]; }
  // output mode t holds input mode perm[t]; stride of output mode t = prod_{u<t} dims[perm[u]]
  int64_t s = 1;
  for (int t = 0; t < nd; ++t) {
    pa.ostride[perm[t]] = s;
    s *= dims[perm[t]];
  }
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_permute, dim3((unsigned)g), dim3(256), 0, st, in, out, total, pa);
}

}  // namespace ttx
