// Device-side helpers shared by every ttx kernel: MFMA operand types, bf16 hi/lo splitting, and
// the counter-based Gaussian generator (Philox4x32-10 + Box-Muller) whose every floating-point
// step is one correctly rounded IEEE binary32 operation (no contraction into FMA), so a CPU
// restatement in numpy float32 reproduces it bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ttx {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Arithmetic used by the streaming GEMM kernels.
//   FP32  : v_mfma_f32_32x32x2_f32 on fp32 operands.
//   BF16  : one v_mfma_f32_32x32x16_bf16 per product on bf16-rounded operands (range finder).
//   BF16X3: operands split x = hi + lo (both bf16); hi*hi + hi*lo + lo*hi with fp32 accumulation,
//           about 16 significant bits per operand (range finder).  Against the bf16-exact Gaussian
//           sketch the unfolding is split three ways instead (hi + mid + lo, three MFMAs), so the
//           sketch and its held-out test columns are accurate to FP32 rounding.
//   FP16  : one v_mfma_f32_32x32x16_f16 per product on fp16-rounded operands.
//   BF16X4: as BF16X3 plus lo*lo: the product error is then the 2^-17 representation error of the
//           split operands (remainder and exact-path Gram).
enum Arith { AR_FP32 = 0, AR_BF16 = 1, AR_BF16X3 = 2, AR_FP16 = 3, AR_BF16X4 = 4 };

// ------------------------------------------------------------------------------------------
// MFMA 32x32 output layout (verified by a GPU test): lane l, accumulator v holds
// row 8*(v/4) + 4*(l/32) + (v%4), column l%32.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int mfma32_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

__device__ __forceinline__ f32x16 mfma_bf16(const bf16x8& a, const bf16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma_f16(const f16x8& a, const f16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma_f32(float a, float b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x16 zero16() {
  f32x16 z;
#pragma unroll
  for (int i = 0; i < 16; ++i) z[i] = 0.0f;
  return z;
}

// Pairs of floats are converted together (v_cvt_pk_bf16_f32, round to nearest even); the float value
// of a packed bf16 is its bit pattern shifted into the upper half, so residuals need no conversion back.
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t pk_bf16(float a, float b) {
  typedef float f32x2v __attribute__((ext_vector_type(2)));
  typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
  const f32x2v v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2v));
}
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

// x = hi + lo with hi = bf16(x) (round to nearest even) and lo = bf16(x - hi)
__device__ __forceinline__ void split8(const float* x, bf16x8& hi, bf16x8& lo) {
  u32x4v h, l;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const float a = x[2 * p], b = x[2 * p + 1];
    const uint32_t w = pk_bf16(a, b);
    h[p] = w;
    l[p] = pk_bf16(a - bf_lo(w), b - bf_hi(w));
  }
  hi = __builtin_bit_cast(bf16x8, h);
  lo = __builtin_bit_cast(bf16x8, l);
}
// x = hi + mid + lo, three bf16 terms (about 24 significant bits): exact products with a bf16-exact
// operand
__device__ __forceinline__ void split8x3(const float* x, bf16x8& hi, bf16x8& mid, bf16x8& lo) {
  u32x4v h, m, l;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const float a = x[2 * p], b = x[2 * p + 1];
    const uint32_t wh = pk_bf16(a, b);
    const float ra = a - bf_lo(wh), rb = b - bf_hi(wh);
    const uint32_t wm = pk_bf16(ra, rb);
    h[p] = wh;
    m[p] = wm;
    l[p] = pk_bf16(ra - bf_lo(wm), rb - bf_hi(wm));
  }
  hi = __builtin_bit_cast(bf16x8, h);
  mid = __builtin_bit_cast(bf16x8, m);
  lo = __builtin_bit_cast(bf16x8, l);
}
__device__ __forceinline__ void round8(const float* x, bf16x8& hi) {
  u32x4v h;
#pragma unroll
  for (int p = 0; p < 4; ++p) h[p] = pk_bf16(x[2 * p], x[2 * p + 1]);
  hi = __builtin_bit_cast(bf16x8, h);
}
__device__ __forceinline__ void round8h(const float* x, f16x8& h) {
#pragma unroll
  for (int i = 0; i < 8; ++i) h[i] = (_Float16)x[i];
}

// ------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller (1958)
// ------------------------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // each 32 x 32 -> 64-bit product as one multiply (v_mad_u64_u32) rather than its low and high
    // halves apart (v_mul_lo_u32 and v_mul_hi_u32)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// ln(u), u in (0,1]: u = 2^e f with f in [sqrt(1/2), sqrt(2)]; ln f = 2 atanh(s), s = (f-1)/(f+1),
// atanh by its odd Taylor series through s^11 (|s| <= 0.1716).
__device__ __forceinline__ float det_log(float u) {
#pragma clang fp contract(off)
  const uint32_t b = __float_as_uint(u);
  int e = (int)(b >> 23) - 127;
  uint32_t mb = (b & 0x7FFFFFu) | 0x3F800000u;
  if (mb > 0x3FB504F3u) { mb -= 0x00800000u; e += 1; }
  const float f = __uint_as_float(mb);
  const float s = __fdiv_rn(f - 1.0f, f + 1.0f);
  const float s2 = s * s;
  float p = 1.0f / 11.0f;
  p = p * s2 + 1.0f / 9.0f;
  p = p * s2 + 1.0f / 7.0f;
  p = p * s2 + 1.0f / 5.0f;
  p = p * s2 + 1.0f / 3.0f;
  p = p * s2;
  const float lnf = (s + s * p) * 2.0f;
  return (float)e * 0.693147180559945309f + lnf;
}

// (cos, sin) of 2 pi (xb >> 8) / 2^24 with exact quadrant reduction and Taylor polynomials
__device__ __forceinline__ void det_sincos(uint32_t xb, float& cs_out, float& sn_out) {
#pragma clang fp contract(off)
  const int w = (int)(xb >> 8);
  const int q = (w + (1 << 21)) >> 22;
  const float d = (float)(w - (q << 22));
  const float phi = d * (float)(6.283185307179586 / 16777216.0);
  const float p2 = phi * phi;
  float hs = -1.0f / 39916800.0f;
  hs = hs * p2 + 1.0f / 362880.0f;
  hs = hs * p2 + -1.0f / 5040.0f;
  hs = hs * p2 + 1.0f / 120.0f;
  hs = hs * p2 + -1.0f / 6.0f;
  const float sn = phi + (phi * p2) * hs;
  float hc = 1.0f / 479001600.0f;
  hc = hc * p2 + -1.0f / 3628800.0f;
  hc = hc * p2 + 1.0f / 40320.0f;
  hc = hc * p2 + -1.0f / 720.0f;
  hc = hc * p2 + 1.0f / 24.0f;
  hc = hc * p2 + -0.5f;
  const float cs = 1.0f + p2 * hc;
  switch (q & 3) {
    case 0: cs_out = cs; sn_out = sn; break;
    case 1: cs_out = -sn; sn_out = cs; break;
    case 2: cs_out = -cs; sn_out = -sn; break;
    default: cs_out = sn; sn_out = -cs; break;
  }
}

__device__ __forceinline__ void box_muller(uint32_t xa, uint32_t xb, float& z0, float& z1) {
#pragma clang fp contract(off)
  const uint32_t t = ((xa >> 9) << 1) | 1u;
  const float u = (float)t * (1.0f / 16777216.0f);
  const float r = sqrtf(-2.0f * det_log(u));   // correctly rounded sqrt (default HIP lowering)
  float c, s;
  det_sincos(xb, c, s);
  z0 = r * c;
  z1 = r * s;
}

// four normals for counter (row_lo, row_hi, group, stream), key (seed_lo, seed_hi)
__device__ __forceinline__ void normal4(uint64_t seed, uint32_t stream, uint64_t row,
                                        uint32_t group, float z[4]) {
  u32x4 c;
  c.x = (uint32_t)row;
  c.y = (uint32_t)(row >> 32);
  c.z = group;
  c.w = stream;
  const u32x4 x = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  box_muller(x.x, x.y, z[0], z[1]);
  box_muller(x.z, x.w, z[2], z[3]);
}

// Sketch entries are rounded to the nearest bfloat16 (ties to even) and |v| < 2^-17 is flushed
// to zero, so every entry is exact in bfloat16, float16 and float32.
__device__ __forceinline__ float omega_round(float v) {
  uint32_t b = __float_as_uint(v);
  b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
  const float r = __uint_as_float(b);
  return (fabsf(r) < 7.62939453125e-06f) ? 0.0f : r;
}

// ------------------------------------------------------------------------------------------
// The block a kernel body works on.  A plain launch takes blockIdx / gridDim; a grouped launch
// (one launch over the same kernel's work for every tensor of a batch) gives each hardware block
// one block of one job: map[blockIdx.x] = {job, block index within that job's grid}, and the job
// descriptor (every descriptor starts with its grid gx x gy) supplies the arguments.  A body run
// on a job therefore does exactly what the plain launch of that job would do.
// ------------------------------------------------------------------------------------------
struct Blk {
  int x, y, gx, gy;
};
__device__ __forceinline__ Blk hw_blk() { return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y}; }

template <class J>
__device__ __forceinline__ const J& grp_job(const J* __restrict__ jobs, const int2* __restrict__ map, Blk& B) {
  const int2 e = map[blockIdx.x];
  const J& j = jobs[e.x];
  B = Blk{e.y % j.gx, e.y / j.gx, j.gx, j.gy};
  return j;
}

}  // namespace ttx
