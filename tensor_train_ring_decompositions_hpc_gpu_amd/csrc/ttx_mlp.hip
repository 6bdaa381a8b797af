// The rank model: a 2 x 128 GELU multilayer perceptron (IN -> 128 -> 128 -> OUT) with hand-written
// forward, backward and Adam kernels.  Every GEMM rounds its weight and activation operands to
// binary16 on load and accumulates in FP32 (the products of two binary16 values are exact in FP32),
// which is the arithmetic of an FP16 MFMA; master weights, gradients and the Adam moments are FP32.
//
// Parameter vector (FP32, row-major blocks): W1 [IN][H], b1 [H], W2 [H][H], b2 [H], W3 [H][OUT], b3 [OUT].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ttx.h"
#include "ttx_device.h"
#include "ttx_kernels.h"

namespace ttx {

__device__ __forceinline__ float r16(float x) { return (float)(_Float16)x; }

// C[m][n] = sum_k opA(m, k) opB(k, n) (+ bias[n]); opA(m, k) = A[m * sam + k * sak] (rounded to
// binary16 when RA), opB(k, n) = B[k * sbk + n * sbn] (rounded when RB).  32 x 32 tile per WG of 256
// threads (2 x 2 outputs each), K staged through LDS in slices of 32.
template <bool RA, bool RB>
__global__ __launch_bounds__(256) void k_mlp_gemm(int M, int N, int K, const float* __restrict__ A, int64_t sam,
                                                  int64_t sak, const float* __restrict__ B, int64_t sbk,
                                                  int64_t sbn, const float* __restrict__ bias,
                                                  float* __restrict__ C, int64_t ldc) {
  __shared__ float as[32][33];
  __shared__ float bs[32][33];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  float acc[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
  for (int k0 = 0; k0 < K; k0 += 32) {
    for (int e = threadIdx.x; e < 32 * 32; e += 256) {
      const int r = e >> 5, c = e & 31;
      const int gm = m0 + r, gk = k0 + c;
      float va = (gm < M && gk < K) ? A[(int64_t)gm * sam + (int64_t)gk * sak] : 0.0f;
      as[r][c] = RA ? r16(va) : va;
      const int gk2 = k0 + r, gn = n0 + c;
      float vb = (gk2 < K && gn < N) ? B[(int64_t)gk2 * sbk + (int64_t)gn * sbn] : 0.0f;
      bs[r][c] = RB ? r16(vb) : vb;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < 32; ++kk) {
      const float a0 = as[ty][kk], a1 = as[ty + 16][kk];
      const float b0 = bs[kk][tx], b1 = bs[kk][tx + 16];
      acc[0][0] += a0 * b0;
      acc[0][1] += a0 * b1;
      acc[1][0] += a1 * b0;
      acc[1][1] += a1 * b1;
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gm = m0 + ty + 16 * i, gn = n0 + tx + 16 * j;
      if (gm < M && gn < N) C[(int64_t)gm * ldc + gn] = acc[i][j] + (bias ? bias[gn] : 0.0f);
    }
}

static void gemm(bool ra, bool rb, int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B,
                 int64_t sbk, int64_t sbn, const float* bias, float* C, int64_t ldc, hipStream_t st) {
  dim3 grid((N + 31) / 32, (M + 31) / 32);
  if (ra && rb) hipLaunchKernelGGL((k_mlp_gemm<true, true>), grid, dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, bias, C, ldc);
  else if (ra) hipLaunchKernelGGL((k_mlp_gemm<true, false>), grid, dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, bias, C, ldc);
  else if (rb) hipLaunchKernelGGL((k_mlp_gemm<false, true>), grid, dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, bias, C, ldc);
  else hipLaunchKernelGGL((k_mlp_gemm<false, false>), grid, dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, bias, C, ldc);
}

// GELU(x) = x Phi(x) = 0.5 x (1 + erf(x / sqrt 2)); GELU'(x) = Phi(x) + x phi(x)
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__global__ void k_gelu(const float* __restrict__ Z, float* __restrict__ H, int64_t count) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
    H[e] = gelu(Z[e]);
}
__global__ void k_gelu_back(const float* __restrict__ dH, const float* __restrict__ Z, float* __restrict__ dZ,
                            int64_t count) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
    dZ[e] = dH[e] * gelu_d(Z[e]);
}

// mean squared error over batch x out: dY = 2 (Y - T) / (B OUT); part[block] = sum (Y - T)^2 (FP64)
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T, int64_t count,
                                             float* __restrict__ dY, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  const float sc = 2.0f / (float)count;
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < count; e += (int64_t)gridDim.x * 256) {
    const float d = Y[e] - T[e];
    dY[e] = sc * d;
    s += (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// db[n] = sum_m dZ[m][n]
__global__ void k_colsum(const float* __restrict__ dZ, int M, int N, float* __restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int m = 0; m < M; ++m) s += dZ[(int64_t)m * N + n];
  db[n] = s;
}

// Adam (Kingma, Ba 2015) with bias correction; t is the 1-based step
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int count, float lr, float beta1, float beta2, float eps, float c1,
                       float c2) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < count; e += gridDim.x * blockDim.x) {
    const float gi = g[e];
    const float a = beta1 * m1[e] + (1.0f - beta1) * gi;
    const float b = beta2 * m2[e] + (1.0f - beta2) * gi * gi;
    m1[e] = a;
    m2[e] = b;
    p[e] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
  }
}

// He-normal initial weights (std sqrt(2 / fan_in)) from the Philox stream, zero biases
__global__ void k_mlp_init(float* __restrict__ p, int in, int hid, int out, uint64_t seed) {
  const int n1 = in * hid, n2 = hid * hid, n3 = hid * out;
  const int o_b1 = n1, o_w2 = o_b1 + hid, o_b2 = o_w2 + n2, o_w3 = o_b2 + hid, o_b3 = o_w3 + n3;
  const int total = o_b3 + out;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    float z[4];
    normal4(seed, 0x4D4Cu, (uint64_t)(e >> 2), 0, z);
    const float zn = z[e & 3];
    float v = 0.0f;
    if (e < o_b1) v = zn * sqrtf(2.0f / (float)in);
    else if (e >= o_w2 && e < o_b2) v = zn * sqrtf(2.0f / (float)hid);
    else if (e >= o_w3 && e < o_b3) v = zn * sqrtf(2.0f / (float)hid);
    p[e] = v;
  }
}

// Features of the rank model from the probe's shifted Cholesky pivots d[i] = R_ii^2 (i < kp):
//   x[b]  = log10(max(d[2b] + d[2b+1], 1e-30 D) / D) / 10 for b < kp / 2 (D = sum d), zero beyond
//   x[32] = log10(rel_tol) / 10,  x[33] = log2(m) / 16,  x[34] = log2(n) / 32,  x[35] = 1, rest 0
__global__ void k_mlp_features(const double* __restrict__ d, int kp, double rel_tol, double m, double n,
                               float* __restrict__ x, int in) {
  __shared__ double D;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < kp; ++i) s += d[i];
    D = s > 0.0 ? s : 1.0;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= in) return;
  double v = 0.0;
  if (t < 32) {
    if (2 * t + 1 < kp) {
      const double pr = fmax(d[2 * t] + d[2 * t + 1], 1e-30 * D);
      v = log10(pr / D) / 10.0;
    }
  } else if (t == 32) {
    v = log10(rel_tol) / 10.0;
  } else if (t == 33) {
    v = log2(m) / 16.0;
  } else if (t == 34) {
    v = log2(n) / 32.0;
  } else if (t == 35) {
    v = 1.0;
  }
  x[t] = (float)v;
}

// One sample end to end in one workgroup (the prediction inside a decomposition step): the
// features of k_mlp_features, then the three layers with the arithmetic of k_mlp_gemm (binary16
// operands, FP32 products accumulated in k order from zero, bias added last) and k_gelu, so the
// output equals that of mlp_features + mlp_forward on the same pivots.  HID threads.
template <int IN, int HID, int OUT>
__global__ __launch_bounds__(HID) void k_mlp_predict(const float* __restrict__ P, const double* __restrict__ d, int kp,
                                                     double rel_tol, double m, double n, float* __restrict__ y) {
  __shared__ float xs[IN];
  __shared__ float h1[HID];
  __shared__ float h2[HID];
  __shared__ double D;
  __shared__ double ds[HID];   // the pivots, loaded once in parallel
  static_assert(HID >= 64, "kp <= 64 pivots (checked by the callers) fit the staging array");
  const float* W1 = P;
  const float* b1 = W1 + IN * HID;
  const float* W2 = b1 + HID;
  const float* b2 = W2 + HID * HID;
  const float* W3 = b2 + HID;
  const float* b3 = W3 + HID * OUT;
  const int t = threadIdx.x;
  if (t < kp) ds[t] = d[t];
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int i = 0; i < kp; ++i) s += ds[i];
    D = s > 0.0 ? s : 1.0;
  }
  __syncthreads();
  if (t < IN) {
    double v = 0.0;
    if (t < 32) {
      if (2 * t + 1 < kp) v = log10(fmax(ds[2 * t] + ds[2 * t + 1], 1e-30 * D) / D) / 10.0;
    } else if (t == 32) {
      v = log10(rel_tol) / 10.0;
    } else if (t == 33) {
      v = log2(m) / 16.0;
    } else if (t == 34) {
      v = log2(n) / 32.0;
    } else if (t == 35) {
      v = 1.0;
    }
    xs[t] = r16((float)v);
  }
  __syncthreads();
  // the layers keep the k-order FP32 chains of mlp_forward; unrolled so the (L2) weight loads of
  // later k are in flight while the chain runs
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < IN; ++k) acc += xs[k] * r16(W1[k * HID + t]);
  h1[t] = r16(gelu(acc + b1[t]));
  __syncthreads();
  acc = 0.0f;
#pragma unroll 32
  for (int k = 0; k < HID; ++k) acc += h1[k] * r16(W2[k * HID + t]);
  h2[t] = r16(gelu(acc + b2[t]));
  __syncthreads();
  if (t < OUT) {
    acc = 0.0f;
#pragma unroll 32
    for (int k = 0; k < HID; ++k) acc += h2[k] * r16(W3[k * OUT + t]);
    y[t] = acc + b3[t];
  }
}

void mlp_predict(const float* P, const double* d, int kp, double rel_tol, double m, double n, float* y,
                 hipStream_t st) {
  hipLaunchKernelGGL((k_mlp_predict<TTX_MLP_IN, TTX_MLP_HIDDEN, TTX_MLP_OUT>), dim3(1), dim3(TTX_MLP_HIDDEN), 0, st, P,
                     d, kp, rel_tol, m, n, y);
}

static unsigned ew_grid(int64_t count) {
  int64_t g = (count + 255) / 256;
  if (g > 2048) g = 2048;
  return (unsigned)(g < 1 ? 1 : g);
}

void mlp_init(float* params, int in, int hid, int out, uint64_t seed, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_init, dim3(64), dim3(256), 0, st, params, in, hid, out, seed);
}

void mlp_features(const double* d, int kp, double rel_tol, double m, double n, float* x, int in, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_features, dim3(1), dim3(64), 0, st, d, kp, rel_tol, m, n, x, in);
}

void mlp_forward(const float* P, int in, int hid, int out, const float* X, int B, float* Z1, float* H1, float* Z2,
                 float* H2, float* Y, hipStream_t st) {
  const float* W1 = P;
  const float* b1 = W1 + in * hid;
  const float* W2 = b1 + hid;
  const float* b2 = W2 + hid * hid;
  const float* W3 = b2 + hid;
  const float* b3 = W3 + hid * out;
  gemm(true, true, B, hid, in, X, in, 1, W1, hid, 1, b1, Z1, hid, st);
  hipLaunchKernelGGL(k_gelu, dim3(ew_grid((int64_t)B * hid)), dim3(256), 0, st, Z1, H1, (int64_t)B * hid);
  gemm(true, true, B, hid, hid, H1, hid, 1, W2, hid, 1, b2, Z2, hid, st);
  hipLaunchKernelGGL(k_gelu, dim3(ew_grid((int64_t)B * hid)), dim3(256), 0, st, Z2, H2, (int64_t)B * hid);
  gemm(true, true, B, out, hid, H2, hid, 1, W3, out, 1, b3, Y, out, st);
}

// gradient of mean((Y - T)^2) into grad (parameter layout); loss partials into part (ew_grid(B out) doubles)
void mlp_backward(const float* P, int in, int hid, int out, const float* X, const float* T, int B, const float* Z1,
                  const float* H1, const float* Z2, const float* H2, const float* Y, float* dY, float* dH,
                  float* dZ, float* grad, double* part, hipStream_t st) {
  const float* W2 = P + in * hid + hid;
  const float* W3 = W2 + hid * hid + hid;
  float* gW1 = grad;
  float* gb1 = gW1 + in * hid;
  float* gW2 = gb1 + hid;
  float* gb2 = gW2 + hid * hid;
  float* gW3 = gb2 + hid;
  float* gb3 = gW3 + hid * out;
  const int64_t cy = (int64_t)B * out, ch = (int64_t)B * hid;
  hipLaunchKernelGGL(k_mse, dim3(ew_grid(cy)), dim3(256), 0, st, Y, T, cy, dY, part);
  // layer 3: gW3 = r16(H2)^T dY, gb3 = colsum dY, dH2 = dY r16(W3)^T
  gemm(true, false, hid, out, B, H2, 1, hid, dY, out, 1, nullptr, gW3, out, st);
  hipLaunchKernelGGL(k_colsum, dim3((out + 127) / 128), dim3(128), 0, st, dY, B, out, gb3);
  gemm(false, true, B, hid, out, dY, out, 1, W3, 1, out, nullptr, dH, hid, st);
  hipLaunchKernelGGL(k_gelu_back, dim3(ew_grid(ch)), dim3(256), 0, st, dH, Z2, dZ, ch);
  // layer 2
  gemm(true, false, hid, hid, B, H1, 1, hid, dZ, hid, 1, nullptr, gW2, hid, st);
  hipLaunchKernelGGL(k_colsum, dim3((hid + 127) / 128), dim3(128), 0, st, dZ, B, hid, gb2);
  gemm(false, true, B, hid, hid, dZ, hid, 1, W2, 1, hid, nullptr, dH, hid, st);
  hipLaunchKernelGGL(k_gelu_back, dim3(ew_grid(ch)), dim3(256), 0, st, dH, Z1, dZ, ch);
  // layer 1
  gemm(true, false, in, hid, B, X, 1, in, dZ, hid, 1, nullptr, gW1, hid, st);
  hipLaunchKernelGGL(k_colsum, dim3((hid + 127) / 128), dim3(128), 0, st, dZ, B, hid, gb1);
}

int mlp_loss_blocks(int64_t count) { return (int)ew_grid(count); }

void mlp_adam(float* P, const float* g, float* m1, float* m2, int count, float lr, int t, hipStream_t st) {
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(beta1, (float)t), c2 = 1.0f - powf(beta2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(ew_grid(count)), dim3(256), 0, st, P, g, m1, m2, count, lr, beta1, beta2, eps, c1,
                     c2);
}

}  // namespace ttx
