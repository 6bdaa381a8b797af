// Streaming kernels: the passes that read a whole unfolding A (m x n, column-major, leading
// dimension lda) from HBM once.
//
//   k_sketch     Y = A X          X = Omega (Philox, generated in-kernel) or X = Z (n x kp, row-major)
//                                 -> per-slab partial sums, reduced by k_reduce_f32
//   k_proj       D = Q^T A        Q column-major (kp x m); D written either as Z = D^T (n x kp,
//                                 row-major: the power pass) or as the next unfolding r x n
//                                 (column-major: the remainder / projection)
//   k_gram_small C = A A^T        m <= 64, exact path (ttdec6_singleGPU.cu:388-401)
//
// Rows of A are spread over lanes so that every global load is a 16-byte vector along a column
// (rows 4r..4r+3 of one column per lane); the MFMA row index r of tile s then maps to physical
// row 4r+s.  The 32x32 MFMA output layout is documented in ttx_device.h.
#include <cstdlib>
#include <type_traits>

#include "ttx_device.h"
#include "ttx_kernels.h"

// cache-policy operand of the load builtins on gfx950: nt (stream past the caches)
constexpr int CPOL_NT = 2;

namespace ttx {

// ------------------------------------------------------------------------------------------------
// sketch: Y = A X
// ------------------------------------------------------------------------------------------------
// WG = 8 waves (2 per SIMD); a wave holds NS = 4 / LT row sub-tiles of 32 rows, so 64 accumulators
// per lane for either sketch width.  RB = row blocks of RW = 32 NS rows per WG, CG = 8 / RB column
// groups.  Wave w: rb = w % RB, cg = w / RB; it accumulates rows [y RW RB + RW rb, +RW) over columns
// jc + 16 cg + [0, 16) of every chunk (chunk width KC = 16 CG); lane r holds rows NS r .. NS r + NS-1.
// Partial output: part[slab][row][kp], slab = blockIdx.x * CG + cg.

template <int LT>
struct SketchAcc {
  f32x16 acc[4 / LT][LT];
};

template <bool VEC, int NS>
__device__ __forceinline__ void sketch_load(const float* __restrict__ A, int64_t lda, int m, int64_t n,
                                            int row, int64_t col0, float (&a)[8][NS]) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int64_t j = col0 + t;
    if (VEC) {
      // full rows: row + NS - 1 < m guaranteed by the caller
      if (j < n) {
        if constexpr (NS == 4) {
          const f32x4 v = *reinterpret_cast<const f32x4*>(A + row + j * lda);
          a[t][0] = v[0]; a[t][1] = v[1]; a[t][2] = v[2]; a[t][3] = v[3];
        } else {
          typedef float f32x2 __attribute__((ext_vector_type(2)));
          const f32x2 v = *reinterpret_cast<const f32x2*>(A + row + j * lda);
          a[t][0] = v[0]; a[t][1] = v[1];
        }
      } else {
#pragma unroll
        for (int s = 0; s < NS; ++s) a[t][s] = 0.0f;
      }
    } else {
#pragma unroll
      for (int s = 0; s < NS; ++s) a[t][s] = (j < n && row + s < m) ? A[row + s + j * lda] : 0.0f;
    }
  }
}

template <int LT, int AR, bool OMEGA>
__device__ __forceinline__ void sketch_compute(const float (&a)[8][4 / LT], const float* __restrict__ xs,
                                               int pitch, int kofs, int lane, SketchAcc<LT>& S, float sa,
                                               const float (&sx)[LT]) {
  constexpr int NS = 4 / LT;
  const int c = lane & 31;
  const int b = lane >> 5;
  // B operands: X[k = kofs + 8b + t][col = 32 lt + c]
  float xv[LT][8];
#pragma unroll
  for (int lt = 0; lt < LT; ++lt) {
    const float* p = xs + (32 * lt + c) * pitch + kofs + 8 * b;
    const f32x4 u0 = *reinterpret_cast<const f32x4*>(p);
    const f32x4 u1 = *reinterpret_cast<const f32x4*>(p + 4);
    xv[lt][0] = u0[0]; xv[lt][1] = u0[1]; xv[lt][2] = u0[2]; xv[lt][3] = u0[3];
    xv[lt][4] = u1[0]; xv[lt][5] = u1[1]; xv[lt][6] = u1[2]; xv[lt][7] = u1[3];
  }
  if constexpr (AR == AR_FP32) {
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int lt = 0; lt < LT; ++lt) S.acc[s][lt] = mfma_f32(a[t][s], xv[lt][t], S.acc[s][lt]);
  } else if constexpr (AR == AR_FP16) {
    // operands scaled by powers of two (A by sa, column c of X by sx) before rounding to fp16
    f16x8 bh[LT];
#pragma unroll
    for (int lt = 0; lt < LT; ++lt) {
#pragma unroll
      for (int t = 0; t < 8; ++t) xv[lt][t] *= sx[lt];
      round8h(xv[lt], bh[lt]);
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = a[t][s] * sa;
      f16x8 ah;
      round8h(v, ah);
#pragma unroll
      for (int lt = 0; lt < LT; ++lt) S.acc[s][lt] = mfma_f16(ah, bh[lt], S.acc[s][lt]);
    }
  } else {
    bf16x8 bh[LT], bl[LT];
#pragma unroll
    for (int lt = 0; lt < LT; ++lt) {
      if (AR == AR_BF16X3 && !OMEGA) split8(xv[lt], bh[lt], bl[lt]);
      else round8(xv[lt], bh[lt]);   // Omega entries are exact in bf16
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = a[t][s];
      bf16x8 ah, am, al;
      if (AR == AR_BF16X3 && OMEGA) split8x3(v, ah, am, al);
      else if (AR == AR_BF16X3) split8(v, ah, al);
      else round8(v, ah);
#pragma unroll
      for (int lt = 0; lt < LT; ++lt) {
        S.acc[s][lt] = mfma_bf16(ah, bh[lt], S.acc[s][lt]);
        if (AR == AR_BF16X3) {
          S.acc[s][lt] = mfma_bf16(al, bh[lt], S.acc[s][lt]);
          if (OMEGA) S.acc[s][lt] = mfma_bf16(am, bh[lt], S.acc[s][lt]);
          else S.acc[s][lt] = mfma_bf16(ah, bl[lt], S.acc[s][lt]);
        }
      }
    }
  }
}

// The LDS chunk xs[c][k] (c < kp = 32 LT, k < KC) holds columns jc .. jc+KC-1 of X.  Its global
// loads (X mode) go to registers one chunk ahead (fill_load) and are written to LDS after the MFMAs
// of the current chunk (fill_store), so neither the A loads nor the MFMAs wait on them; with the
// in-kernel sketch, fill_store generates the Gaussian entries instead.  A thread handles FQ float4
// pieces per chunk (KC kp / 4 / threads = kp / (16 RB), at least 1).
template <bool OMEGA, int FQ>
struct FillRegs {
  f32x4 v[OMEGA ? 1 : FQ];
};

template <int LT, bool OMEGA, int FQ>
__device__ __forceinline__ void fill_load(FillRegs<OMEGA, FQ>& R, int KC, int64_t jc, int64_t n,
                                          const float* __restrict__ X, int64_t ldx) {
  if constexpr (!OMEGA) {
    constexpr int NG = 32 * LT / 4;
    const int total = KC * NG;
#pragma unroll
    for (int q = 0; q < FQ; ++q) {
      const int idx = threadIdx.x + q * (int)blockDim.x;
      if (idx < total) {
        const int jj = idx % KC;
        const int g = idx / KC;
        const int64_t j = jc + jj;
        if (j < n) R.v[q] = *reinterpret_cast<const f32x4*>(X + j * ldx + 4 * g);
        else R.v[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      }
    }
  }
}

template <int LT, bool OMEGA, int FQ>
__device__ __forceinline__ void fill_store(const FillRegs<OMEGA, FQ>& R, float* __restrict__ xs, int pitch, int KC,
                                           int64_t jc, int64_t n, uint64_t seed, uint32_t stream, int ell) {
  constexpr int NG = 32 * LT / 4;
  const int total = KC * NG;
#pragma unroll
  for (int q = 0; q < (OMEGA ? 4 : FQ); ++q) {
    const int idx = threadIdx.x + q * (int)blockDim.x;
    if (idx < total) {
      const int jj = idx % KC;
      const int g = idx / KC;
      const int64_t j = jc + jj;
      float z[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if constexpr (OMEGA) {
        if (j < n && 4 * g < ell) {
          normal4(seed, stream, (uint64_t)j, (uint32_t)g, z);
#pragma unroll
          for (int e = 0; e < 4; ++e) z[e] = (4 * g + e < ell) ? omega_round(z[e]) : 0.0f;
        }
      } else {
        z[0] = R.v[q][0]; z[1] = R.v[q][1]; z[2] = R.v[q][2]; z[3] = R.v[q][3];
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) xs[(4 * g + e) * pitch + jj] = z[e];
    }
  }
}

// One-shot fill of the LDS chunk (loads, or Gaussian entries, written straight to LDS): the order of
// the variants whose registers do not hold a chunk of X across the MFMAs.
template <int LT, bool OMEGA>
__device__ __forceinline__ void sketch_fill(float* __restrict__ xs, int pitch, int KC, int64_t jc, int64_t n,
                                            const float* __restrict__ X, int64_t ldx,
                                            uint64_t seed, uint32_t stream, int ell) {
  constexpr int NG = 32 * LT / 4;
  const int total = KC * NG;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int jj = idx % KC;
    const int g = idx / KC;
    const int64_t j = jc + jj;
    float z[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (j < n) {
      if (OMEGA) {
        if (4 * g < ell) {
          normal4(seed, stream, (uint64_t)j, (uint32_t)g, z);
#pragma unroll
          for (int e = 0; e < 4; ++e) z[e] = (4 * g + e < ell) ? omega_round(z[e]) : 0.0f;
        }
      } else {
        const f32x4 v = *reinterpret_cast<const f32x4*>(X + j * ldx + 4 * g);
        z[0] = v[0]; z[1] = v[1]; z[2] = v[2]; z[3] = v[3];
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) xs[(4 * g + e) * pitch + jj] = z[e];
  }
}

template <int LT, int AR, bool OMEGA, int FQ>
__device__ __forceinline__ void sketch_body(const float* __restrict__ A, int64_t lda, int m, int64_t n,
                                            const float* __restrict__ X, int64_t ldx, uint64_t seed,
                                            uint32_t stream, int ell, int RB, int CG, int64_t chunks_per_wg,
                                            float* __restrict__ part, const float* __restrict__ scales, const Blk B) {
  constexpr int KP = 32 * LT;
  const int KC = 16 * CG;
  const int pitch = KC + 4;
  extern __shared__ __align__(16) float smem[];
  float* xs0 = smem;
  float* xs1 = smem + KP * pitch;

  constexpr int NS = 4 / LT;
  constexpr int RW = 32 * NS;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int rb = w % RB;
  const int cg = w / RB;
  const int r = lane & 31;
  const int b = lane >> 5;
  const int row = B.y * RW * RB + RW * rb + NS * r;
  const bool vec = ((lda & (NS - 1)) == 0) && (row + NS - 1 < m);

  const int64_t nchunks = (n + KC - 1) / KC;
  const int64_t c0 = (int64_t)B.x * chunks_per_wg;
  const int64_t c1 = (c0 + chunks_per_wg < nchunks) ? c0 + chunks_per_wg : nchunks;

  SketchAcc<LT> S;
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int lt = 0; lt < LT; ++lt) S.acc[s][lt] = zero16();

  // FP16: power-of-two operand scales (scales[0] for A, scales[1 + c] for column c of X)
  float sa = 1.0f, sx[LT];
#pragma unroll
  for (int lt = 0; lt < LT; ++lt) sx[lt] = 1.0f;
  if (AR == AR_FP16 && scales) {
    sa = scales[0];
    if (!OMEGA)
#pragma unroll
      for (int lt = 0; lt < LT; ++lt) sx[lt] = scales[1 + 32 * lt + r];
  }

  float a0[8][NS], a1[8][NS];
  [[maybe_unused]] FillRegs<OMEGA, FQ> R;
  const int kofs = 16 * cg;
  auto loadA = [&](float (&a)[8][NS], int64_t c) {
    const int64_t col = c * KC + kofs + 8 * b;
    if (vec) sketch_load<true, NS>(A, lda, m, n, row, col, a);
    else sketch_load<false, NS>(A, lda, m, n, row, col, a);
  };
  // the pipelined order (X chunk loads held in registers across the MFMAs) fits the registers of the
  // 64-column sketch with at most two pieces per thread; the others fill before computing
  constexpr bool PIPE = (LT == 2) && (FQ <= 2);
  if constexpr (PIPE) {
    if (c0 < c1) {
      fill_load<LT, OMEGA, FQ>(R, KC, c0 * KC, n, X, ldx);
      loadA(a0, c0);
      fill_store<LT, OMEGA, FQ>(R, xs0, pitch, KC, c0 * KC, n, seed, stream, ell);
    }
    __syncthreads();
    for (int64_t c = c0; c < c1; c += 2) {
      // even step: chunk c from a0 / xs0; chunk c+1's loads go out first, its LDS fill after the MFMAs
      if (c + 1 < c1) {
        loadA(a1, c + 1);
        fill_load<LT, OMEGA, FQ>(R, KC, (c + 1) * KC, n, X, ldx);
      }
      sketch_compute<LT, AR, OMEGA>(a0, xs0, pitch, kofs, lane, S, sa, sx);
      __builtin_amdgcn_sched_barrier(0);
      if (c + 1 < c1) fill_store<LT, OMEGA, FQ>(R, xs1, pitch, KC, (c + 1) * KC, n, seed, stream, ell);
      __syncthreads();
      if (c + 1 >= c1) break;
      // odd step: chunk c+1 from a1 / xs1
      if (c + 2 < c1) {
        loadA(a0, c + 2);
        fill_load<LT, OMEGA, FQ>(R, KC, (c + 2) * KC, n, X, ldx);
      }
      sketch_compute<LT, AR, OMEGA>(a1, xs1, pitch, kofs, lane, S, sa, sx);
      __builtin_amdgcn_sched_barrier(0);
      if (c + 2 < c1) fill_store<LT, OMEGA, FQ>(R, xs0, pitch, KC, (c + 2) * KC, n, seed, stream, ell);
      __syncthreads();
    }
  } else {
    if (c0 < c1) {
      sketch_fill<LT, OMEGA>(xs0, pitch, KC, c0 * KC, n, X, ldx, seed, stream, ell);
      loadA(a0, c0);
    }
    __syncthreads();
    for (int64_t c = c0; c < c1; c += 2) {
      if (c + 1 < c1) {
        sketch_fill<LT, OMEGA>(xs1, pitch, KC, (c + 1) * KC, n, X, ldx, seed, stream, ell);
        loadA(a1, c + 1);
      }
      sketch_compute<LT, AR, OMEGA>(a0, xs0, pitch, kofs, lane, S, sa, sx);
      __syncthreads();
      if (c + 1 >= c1) break;
      if (c + 2 < c1) {
        sketch_fill<LT, OMEGA>(xs0, pitch, KC, (c + 2) * KC, n, X, ldx, seed, stream, ell);
        loadA(a0, c + 2);
      }
      sketch_compute<LT, AR, OMEGA>(a1, xs1, pitch, kofs, lane, S, sa, sx);
      __syncthreads();
    }
  }

  // partial output (row-major, kp columns), the FP16 operand scales undone (exact powers of two)
  const int64_t slab = (int64_t)B.x * CG + cg;
  float* out = part + slab * (int64_t)m * KP;
  float unscale[LT];
#pragma unroll
  for (int lt = 0; lt < LT; ++lt) unscale[lt] = 1.0f / (sa * sx[lt]);
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int lt = 0; lt < LT; ++lt)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int prow = B.y * RW * RB + RW * rb + NS * mfma32_row(v, lane) + s;
        if (prow < m) out[(int64_t)prow * KP + 32 * lt + r] = S.acc[s][lt][v] * unscale[lt];
      }
}

template <int LT, int AR, bool OMEGA, int FQ>
__global__ __launch_bounds__(512) void k_sketch(const float* __restrict__ A, int64_t lda, int m, int64_t n,
                                                const float* __restrict__ X, int64_t ldx, uint64_t seed,
                                                uint32_t stream, int ell, int RB, int CG, int64_t chunks_per_wg,
                                                float* __restrict__ part, const float* __restrict__ scales) {
  sketch_body<LT, AR, OMEGA, FQ>(A, lda, m, n, X, ldx, seed, stream, ell, RB, CG, chunks_per_wg, part, scales, hw_blk());
}
template <int LT, int AR, bool OMEGA, int FQ>
__global__ __launch_bounds__(512) void k_sketch_grp(const JobSketch* __restrict__ jobs, const int2* __restrict__ map) {
  Blk B;
  const JobSketch& j = grp_job(jobs, map, B);
  sketch_body<LT, AR, OMEGA, FQ>(j.A, j.lda, j.m, j.n, j.X, j.ldx, j.seed, j.stream, j.ell, j.RB, j.CG, j.cpw, j.part,
                                 nullptr, B);
}

// ------------------------------------------------------------------------------------------------
// sketch with the in-kernel Omega, A streamed straight into LDS
// ------------------------------------------------------------------------------------------------
// Same tiling, arithmetic and output as k_sketch<LT, AR, true, *>, but each wave's A tile (16
// columns x RW rows per chunk) travels from HBM to LDS by global_load_lds_dwordx4, ST chunks deep:
// the bytes in flight no longer cost registers, so a CU keeps (ST - 1) x 32 KB of A requested while
// it computes (a register double buffer allowed one chunk).  Every wave loads and reads only its
// own tile, so the A stages need no barrier: s_waitcnt vmcnt((ST - 1) NI) before reading chunk c
// leaves the ST - 1 later chunks in flight (no other vector-memory operation runs in the loop).
// Load layout: instruction q of a chunk writes 1 KB = CPI = 256 / RW columns of RW rows (then PAD
// floats of padding); lane l reads rows 4 (l % LPC) .. +3 of column CPI q + l / LPC (LPC = RW / 4
// lanes per column), so the stage is the tile in column-major order and lane r of the MFMA reads
// its NS rows of column 8 b + t as one ds_read.  Requires lda % 4 == 0 and m % 4 == 0 (16-byte
// rows); rows past m load row 0 of the column instead (their outputs are not written) and columns
// past n load column n - 1 (zeroed before use).
//
// GRAM0 (LT = 2, BF16X3, m % 64 == 0): the same pass also accumulates the Gram C = A0 A0^T of the
// 32-row view A0 (32 x 32 n) of the same memory: a 64-row tile holds two 32-row columns of A0 per
// column of A.  It is the exact first step of a TT sweep whose second step this sketch serves, so
// the tensor is read once for both.  MFMA row r of sketch tile s is physical row 32 s + r, so the
// tile's A operand is also the Gram operand of half s: the hi and mid terms of the sketch's
// three-way split are the hi / lo of k_gram_small<1, AR_BF16X4>'s two-way split, and the four
// products hi hi, lo hi, hi lo, lo lo go to one FP32 accumulator per half (two interleaved chains),
// flushed to FP64 every 32 chunks; one FP64 32 x 32 partial per WG in
// gpart[blockIdx.y * gridDim.x + blockIdx.x].  G3: the Gram without the mid mid product (three bf16 products, about
// 2^-16 relative per term instead of 2^-17).  PAD = 16 puts the two lane halves' column reads on
// disjoint banks.
// Omega for the precomputed-Omega first pass: chunk c holds rows 16 c .. 16 c + 15 of Omega (the
// Philox counters j), column-major inside the chunk: entry (col, jj) at out[c * 1024 + col * 16 + jj],
// so a chunk is 2 KB and a column's 16 entries are 32 contiguous bytes.  The entries are exactly the
// ones fill_store writes for kp = 64 (omega_round of the normals, exact in bf16; zero for col >= ell
// and for j >= n).  Thread: one chunk and one group of four columns.
__global__ __launch_bounds__(256) void k_omega_chunks(uint64_t seed, uint32_t stream, int64_t n, int ell,
                                                      int64_t nchunks, __bf16* __restrict__ out) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nchunks * 16; t += (int64_t)gridDim.x * 256) {
    const int64_t c = t >> 4;
    const int g = (int)(t & 15);
    bf16x8 v[4][2];
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) {
      const int64_t j = c * 16 + jj;
      float z[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (j < n && 4 * g < ell) {
        normal4(seed, stream, (uint64_t)j, (uint32_t)g, z);
#pragma unroll
        for (int e = 0; e < 4; ++e) z[e] = (4 * g + e < ell) ? omega_round(z[e]) : 0.0f;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e][jj >> 3][jj & 7] = (__bf16)z[e];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      bf16x8* o = reinterpret_cast<bf16x8*>(out + c * 1024 + (4 * g + e) * 16);
      o[0] = v[e][0];
      o[1] = v[e][1];
    }
  }
}

template <int LT, int AR, int ST, bool GRAM0, bool OP = false, bool G3 = false>
__global__ __launch_bounds__(512) void k_sketch_dl(const float* __restrict__ A, int64_t lda, int m, int64_t n,
                                                   uint64_t seed, uint32_t stream, int ell, int RB, int CG,
                                                   int64_t chunks_per_wg, float* __restrict__ part,
                                                   const float* __restrict__ scales, double* __restrict__ gpart,
                                                   const __bf16* __restrict__ omega = nullptr) {
  constexpr int KP = 32 * LT;
  constexpr int NS = 4 / LT;
  constexpr int RW = 32 * NS;
  constexpr int LPC = RW / 4;            // lanes per column in a load
  constexpr int CPI = 64 / LPC;          // columns per load instruction
  constexpr int NI = 16 / CPI;           // load instructions per chunk and wave
  constexpr int PAD = GRAM0 ? 16 : 0;
  constexpr int TILE = NI * (256 + PAD); // floats of a wave's tile
  static_assert(!GRAM0 || (LT == 2 && AR == AR_BF16X3), "the fused Gram takes 64-row tiles, BF16X3");
  // OP: Omega comes precomputed (k_omega_chunks layout, bf16) through a ring of ST + 1 chunk slots
  // of 2 KB in the space of the two generated chunks; waves 0 and 1 load a chunk's halves
  // direct-to-LDS with their A tiles
  static_assert(!OP || (GRAM0 && (ST + 1) * 2048 <= 2 * KP * 20 * 4), "the Omega ring fits the chunk space");
  const int KC = 16 * CG;
  const int pitch = KC + 4;
  extern __shared__ __align__(16) float smem[];
  float* xs0 = smem;
  float* xs1 = smem + KP * pitch;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  float* stg = smem + 2 * KP * pitch + w * TILE;   // stage s of this wave at stg + s * nw * TILE

  const int rb = w % RB;
  const int cg = w / RB;
  const int r = lane & 31;
  const int b = lane >> 5;
  const int row0 = blockIdx.y * RW * RB + RW * rb;

  const int64_t nchunks = (n + KC - 1) / KC;
  const int64_t c0 = (int64_t)blockIdx.x * chunks_per_wg;
  const int64_t c1 = (c0 + chunks_per_wg < nchunks) ? c0 + chunks_per_wg : nchunks;

  SketchAcc<LT> S;
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int lt = 0; lt < LT; ++lt) S.acc[s][lt] = zero16();
  float sa = 1.0f, sx[LT];
#pragma unroll
  for (int lt = 0; lt < LT; ++lt) sx[lt] = 1.0f;
  if (AR == AR_FP16 && scales) sa = scales[0];
  [[maybe_unused]] f32x16 gacc[2];   // one per half: two independent MFMA chains
  [[maybe_unused]] double gd[16];
  if constexpr (GRAM0) {
    gacc[0] = zero16();
    gacc[1] = zero16();
#pragma unroll
    for (int v = 0; v < 16; ++v) gd[v] = 0.0;
  }

  // column col of the tile starts at tile_col(col) floats into the stage
  auto tile_col = [](int col) { return (col / CPI) * (256 + PAD) + (col % CPI) * RW; };
  // this lane's load: rows lrow .. lrow + 3 (row 0 when past m) of column lcol + CPI q of the chunk
  const int lrow_raw = row0 + 4 * (lane % LPC);
  const int lrow = (lrow_raw < m) ? lrow_raw : 0;
  const int lcol = 16 * cg + lane / LPC;
  __attribute__((address_space(3))) char* om_lds = (__attribute__((address_space(3))) char*)smem;
  auto issue = [&](int64_t c) {
    if constexpr (OP) {
      if (w < 2)
        __builtin_amdgcn_global_load_lds((const void*)(omega + c * 1024 + w * 512 + lane * 8),
                                         (__attribute__((address_space(3))) void*)(om_lds + (int)(c % (ST + 1)) * 2048 +
                                                                                   w * 1024),
                                         16, 0, 0);
    }
    float* dst = stg + (int)(c % ST) * nw * TILE;
#pragma unroll
    for (int q = 0; q < NI; ++q) {
      int64_t j = c * KC + lcol + CPI * q;
      if (j >= n) j = n - 1;
      // nontemporal (cache policy nt): the tensor is streamed once per pass (measured 0.89 -> 0.84 ms for
      // the 4.3 GB first pass)
      __builtin_amdgcn_global_load_lds((const void*)(A + j * lda + lrow),
                                       (__attribute__((address_space(3))) void*)(dst + q * (256 + PAD)), 16, 0, CPOL_NT);
    }
  };
  auto compute = [&](int64_t c, const float* xs, int kofs) {
    const float* src = stg + (int)(c % ST) * nw * TILE;
    const bool tail = c * KC + KC > n;
    float a[8][NS];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const float* pp = src + tile_col(8 * b + t);
      if constexpr (GRAM0) {
        // MFMA row r of tile s is physical row 32 s + r: tile s is then also the Gram operand of
        // the 32-row half s, with the same lanes and the same split
        a[t][0] = pp[r];
        a[t][1] = pp[32 + r];
      } else if constexpr (NS == 4) {
        const f32x4 v = *reinterpret_cast<const f32x4*>(pp + NS * r);
        a[t][0] = v[0]; a[t][1] = v[1]; a[t][2] = v[2]; a[t][3] = v[3];
      } else {
        typedef float f32x2 __attribute__((ext_vector_type(2)));
        const f32x2 v = *reinterpret_cast<const f32x2*>(pp + NS * r);
        a[t][0] = v[0]; a[t][1] = v[1];
      }
    }
    if (tail) {
      // tail chunk: columns past n hold a copy of column n - 1
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (c * KC + 16 * cg + 8 * b + t >= n)
#pragma unroll
          for (int s2 = 0; s2 < NS; ++s2) a[t][s2] = 0.0f;
    }
    if constexpr (GRAM0) {
      // sketch_compute<2, AR_BF16X3, true> with the Gram products on the same split operands
      const int cc = lane & 31;
      bf16x8 bh[LT];
#pragma unroll
      for (int lt = 0; lt < LT; ++lt) {
        if constexpr (OP) {
          const __bf16* om = reinterpret_cast<const __bf16*>(smem) + (int)(c % (ST + 1)) * 1024;
          bh[lt] = *reinterpret_cast<const bf16x8*>(om + (32 * lt + cc) * 16 + 8 * b);
        } else {
          const float* p = xs + (32 * lt + cc) * pitch + kofs + 8 * b;
          const f32x4 u0 = *reinterpret_cast<const f32x4*>(p);
          const f32x4 u1 = *reinterpret_cast<const f32x4*>(p + 4);
          const float xv[8] = {u0[0], u0[1], u0[2], u0[3], u1[0], u1[1], u1[2], u1[3]};
          round8(xv, bh[lt]);   // Omega entries are exact in bf16
        }
      }
      const bool gram_rows_valid = row0 < m;   // a tile past m (m % 64 == 0) holds row-0 copies
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) {
        float v[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) v[t] = a[t][s2];
        bf16x8 ah, am, al;
        split8x3(v, ah, am, al);
#pragma unroll
        for (int lt = 0; lt < LT; ++lt) {
          S.acc[s2][lt] = mfma_bf16(ah, bh[lt], S.acc[s2][lt]);
          S.acc[s2][lt] = mfma_bf16(al, bh[lt], S.acc[s2][lt]);
          S.acc[s2][lt] = mfma_bf16(am, bh[lt], S.acc[s2][lt]);
        }
        // Gram of the half s2: hi / lo of the two-term split are hi / mid of the three-term one
        if (gram_rows_valid) {
          gacc[s2] = mfma_bf16(ah, ah, gacc[s2]);
          gacc[s2] = mfma_bf16(am, ah, gacc[s2]);
          gacc[s2] = mfma_bf16(ah, am, gacc[s2]);
          if constexpr (!G3) gacc[s2] = mfma_bf16(am, am, gacc[s2]);
        }
      }
    } else {
      sketch_compute<LT, AR, true>(a, xs, pitch, kofs, lane, S, sa, sx);
    }
  };
  constexpr int WAIT = (ST - 1) * NI;
  constexpr int WAIT_IMM = (WAIT & 15) | ((WAIT >> 4) << 14) | 0x0F70;   // vmcnt(WAIT), others unchanged
  static_assert(WAIT < 64, "vmcnt range");
  // workgroup barrier for the Omega chunks only: their LDS writes complete (lgkmcnt(0)), then
  // s_barrier.  __syncthreads would also wait for the A loads still in flight (vmcnt(0)), which
  // would drain the pipeline every chunk.
  auto xs_barrier = [] {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0); vmcnt, expcnt unchanged
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };
  if constexpr (OP) {
    // waves 0 and 1 issue one more load per chunk (their Omega halves, ahead of the A tiles)
    constexpr int WAIT_O = (ST - 1) * (NI + 1);
    constexpr int WAIT_O_IMM = (WAIT_O & 15) | ((WAIT_O >> 4) << 14) | 0x0F70;
    static_assert(WAIT_O < 64, "vmcnt range");
    for (int d = 0; d < ST - 1; ++d)
      if (c0 + d < c1) issue(c0 + d);
    for (int64_t c = c0; c < c1; ++c) {
      if (c + ST - 1 < c1) {
        issue(c + ST - 1);
        if (w < 2) __builtin_amdgcn_s_waitcnt(WAIT_O_IMM);
        else __builtin_amdgcn_s_waitcnt(WAIT_IMM);
      } else {
        __builtin_amdgcn_s_waitcnt(0x0F70);   // the last chunks: wait for all
      }
      // chunk c's Omega (loaded by waves 0 and 1) is in LDS for every wave; every wave is also past
      // compute(c - 1), so the slot that issue(c + ST) overwrites next is free
      xs_barrier();
      compute(c, nullptr, 16 * cg);
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (GRAM0) {
        if (((c - c0) & 31) == 31) {
#pragma unroll
          for (int v = 0; v < 16; ++v) gd[v] += (double)gacc[0][v] + (double)gacc[1][v];
          gacc[0] = zero16();
          gacc[1] = zero16();
        }
      }
    }
  } else {
  FillRegs<true, 1> R;
  for (int d = 0; d < ST - 1; ++d)
    if (c0 + d < c1) issue(c0 + d);
  if (c0 < c1) fill_store<LT, true, 1>(R, xs0, pitch, KC, c0 * KC, n, seed, stream, ell);
  xs_barrier();
  for (int64_t c = c0; c < c1; ++c) {
    const float* xcur = ((c - c0) & 1) ? xs1 : xs0;
    float* xnext = ((c - c0) & 1) ? xs0 : xs1;
    // chunk c + ST - 1 goes out before chunk c is read: ST chunks requested; the wait leaves ST - 1
    if (c + ST - 1 < c1) issue(c + ST - 1);
    // the next chunk's Omega entries are generated while this chunk's loads are still in flight
    // (its buffer was last read before the previous barrier); measured 8% faster than after the MFMAs
    if (c + 1 < c1) fill_store<LT, true, 1>(R, xnext, pitch, KC, (c + 1) * KC, n, seed, stream, ell);
    __builtin_amdgcn_sched_barrier(0);
    if (c + ST - 1 < c1) {
      __builtin_amdgcn_s_waitcnt(WAIT_IMM);
    } else {
      __builtin_amdgcn_s_waitcnt(0x0F70);   // the last chunks: wait for all
    }
    compute(c, xcur, 16 * cg);
    __builtin_amdgcn_sched_barrier(0);
    xs_barrier();
    if constexpr (GRAM0) {
      if (((c - c0) & 31) == 31) {
#pragma unroll
        for (int v = 0; v < 16; ++v) gd[v] += (double)gacc[0][v] + (double)gacc[1][v];
        gacc[0] = zero16();
        gacc[1] = zero16();
      }
    }
  }

  }
  const int64_t slab = (int64_t)blockIdx.x * CG + cg;
  float* out = part + slab * (int64_t)m * KP;
  float unscale[LT];
#pragma unroll
  for (int lt = 0; lt < LT; ++lt) unscale[lt] = 1.0f / (sa * sx[lt]);
#pragma unroll
  for (int s2 = 0; s2 < NS; ++s2)
#pragma unroll
    for (int lt = 0; lt < LT; ++lt)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int prow = GRAM0 ? row0 + 32 * s2 + mfma32_row(v, lane) : row0 + NS * mfma32_row(v, lane) + s2;
        if (prow < m) out[(int64_t)prow * KP + 32 * lt + r] = S.acc[s2][lt][v] * unscale[lt];
      }
  if constexpr (GRAM0) {
#pragma unroll
    for (int v = 0; v < 16; ++v) gd[v] += (double)gacc[0][v] + (double)gacc[1][v];
    // the waves' partials through the (now idle) stage area, summed in wave order
    __syncthreads();
    double* red = reinterpret_cast<double*>(smem + 2 * KP * pitch);
#pragma unroll
    for (int v = 0; v < 16; ++v) red[w * 1024 + mfma32_row(v, lane) * 32 + r] = gd[v];
    __syncthreads();
    double* o = gpart + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 1024;
    for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
      double sum = 0.0;
      for (int ww = 0; ww < nw; ++ww) sum += red[ww * 1024 + e];
      o[e] = sum;
    }
  }
}

// ------------------------------------------------------------------------------------------------
// projection: D = Q^T A (kp x n), Q column-major kp x m (Qc[c*ldq + i])
// ------------------------------------------------------------------------------------------------
// WG = 8 waves; wave w owns columns jb + 32 w + [0, 32) of the current column block (256 columns).
// K loop in chunks of 64 rows, starting at chunk blockIdx.x mod nk; the Q chunk (64 rows x kp) is
// staged in LDS and shared by the waves.
// Lane (jl = lane%32, b = lane/32) loads rows 64 kc + 32 b + [0, 32) of column jb + 32 w + jl
// (one 128-byte line); K-step t uses rows 32 b + 8 t + [0, 8) of the chunk for both operands.
// GRAM: also accumulate the Gram D D^T of the written tile (the projected Gram B B^T of the final
// pass): each wave transposes its tile through LDS and runs FP32 MFMAs over its columns; at the end
// the 8 wave sums are added in wave order in FP64 and gpart[blockIdx.x][kg][kg] is written.
// COLMAX (mode 0): also max_j |Z[j][c]| per column c as float bits, atomicMax into colmax[c] (the FP16
// range finder's column scales of Z, without a pass of their own)
template <int LT, int AR, bool GRAM, bool COLMAX = false>
__device__ __forceinline__ void proj_body(const float* __restrict__ A, int64_t lda, int m, int64_t n,
                                          const float* __restrict__ Qc, int64_t ldq,
                                          int out_mode, float* __restrict__ out, int64_t ldo, int r_out,
                                          int64_t blocks_total, double* __restrict__ gpart, int kg,
                                          const float* __restrict__ scales, const Blk B,
                                          unsigned* __restrict__ colmax = nullptr) {
  constexpr int KP = 32 * LT;
  constexpr int QP = 64 + 8;   // LDS pitch of a Q column chunk (floats; bf16 halves use the same pitch)
  // bf16 arithmetic keeps the Q chunk in LDS already split: hi in the first half of a buffer, lo in
  // the second (pitch QP bf16), so no wave re-splits Q
  constexpr bool QSPLIT = (AR == AR_BF16 || AR == AR_BF16X3 || AR == AR_BF16X4);
  // FP16: A is scaled by the power of two scales[0] before rounding, the products are unscaled
  const float sa = (AR == AR_FP16 && scales) ? scales[0] : 1.0f;
  constexpr int NPG = LT * (LT + 1) / 2;
  constexpr int SP = 64 + 4;   // LDS pitch of a staged A column chunk
  __shared__ float qs[2][KP * QP];
  // per-wave staging of its A tile (32 columns x 64 rows); the Gram tile reuses it after the K loop
  __shared__ float stg[8][32 * SP];
  __shared__ double gs[GRAM ? KP * KP : 1];
  __shared__ unsigned cmx[COLMAX ? KP : 1];
  // COLMAX: running max |D[c][j]| of the lane's rows c over its columns, reduced across lanes at the end
  [[maybe_unused]] float amax[COLMAX ? LT : 1][16];
  if constexpr (COLMAX) {
    if (threadIdx.x < KP) cmx[threadIdx.x] = 0u;
#pragma unroll
    for (int lt = 0; lt < LT; ++lt)
#pragma unroll
      for (int v = 0; v < 16; ++v) amax[lt][v] = 0.0f;
  }
  f32x16 gacc[GRAM ? NPG : 1];
  if constexpr (GRAM) {
#pragma unroll
    for (int p = 0; p < NPG; ++p) gacc[p] = zero16();
  }
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int jl = lane & 31;
  const int b = lane >> 5;
  const int nk = (m + 63) / 64;
  // each WG runs its K loop from chunk (its block index) mod nk (wrapping): at a power-of-two column stride
  // every WG reading the same chunk index at once puts all requests on the same address bits below
  // the stride, i.e. on a few DRAM channels; staggered starts spread them (measured 1.13 -> 0.93 ms
  // for the 1024 x 1M step)
  const int kofs = (int)(B.x % nk);
  const bool vec = ((lda & 3) == 0) && ((m & 3) == 0);
  // whole 64-row chunks and 16-byte rows of A and Q: no per-load row guards
  const bool rows_whole = vec && ((m & 63) == 0) && ((ldq & 3) == 0);

  // Q chunk kc (64 rows x KP columns): QF float4 pieces per thread, loaded to registers one chunk
  // ahead (qload) and written to LDS after the MFMAs of the current chunk (qstore), so the A loads
  // and the MFMAs never wait on these (L2) loads
  constexpr int QF = KP * 16 / 512;
  auto qload = [&](f32x4 (&qr)[QF], int kc, auto whole) {
#pragma unroll
    for (int u = 0; u < QF; ++u) {
      const int idx = threadIdx.x + 512 * u;
      const int c = idx / 16;
      const int i = ((kc + kofs) % nk) * 64 + (idx % 16) * 4;
      f32x4 v;
      if constexpr (decltype(whole)::value) {
        // Q's kp columns from one descriptor: the lane's fixed offset plus the chunk's scalar offset
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void*)Qc, 0, (int)(KP * ldq * 4), 0x00020000);
        v = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(
                                          rs, (int)((c * ldq + (idx % 16) * 4) * 4), ((kc + kofs) % nk) * 256, 0));
      } else if (((ldq & 3) == 0) && i + 3 < m) {
        v = *reinterpret_cast<const f32x4*>(Qc + (int64_t)c * ldq + i);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = (i + e < m) ? Qc[(int64_t)c * ldq + i + e] : 0.0f;
      }
      qr[u] = v;
    }
  };
  auto qstore = [&](const f32x4 (&qr)[QF], float* q) {
#pragma unroll
    for (int u = 0; u < QF; ++u) {
      const int idx = threadIdx.x + 512 * u;
      if constexpr (QSPLIT) {
        __bf16* qh = reinterpret_cast<__bf16*>(q);
        __bf16* ql = qh + KP * QP;
        typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
        bf16x4 h, l;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const __bf16 hv = (__bf16)qr[u][e];
          h[e] = hv;
          l[e] = (__bf16)(qr[u][e] - (float)hv);
        }
        const int o = (idx / 16) * QP + (idx % 16) * 4;
        *reinterpret_cast<bf16x4*>(qh + o) = h;
        *reinterpret_cast<bf16x4*>(ql + o) = l;
      } else {
        *reinterpret_cast<f32x4*>(q + (idx / 16) * QP + (idx % 16) * 4) = qr[u];
      }
    }
  };
  // coalesced loads of the wave's tile: load u, lane l takes rows 4 (l % 16) .. +3 of tile column
  // 4 u + l / 16 (16 lanes read one column's 256 bytes); the tile is transposed to the MFMA layout
  // through the wave's LDS staging area
  const int lc = lane >> 4, lr = (lane & 15) * 4;
  // an interior block (all 256 columns below n, whole 64-row chunks): no guards.  Buffer loads from
  // a descriptor at the block's first column: the lane's 32-bit offset never changes and the chunk
  // and load offsets are scalars, so the lane holds one address register instead of a 64-bit pair
  // per load (with the pairs, register pressure put addresses in registers that loads in flight
  // still had to write, and the next chunk's loads waited for this chunk's)
  const int aoff = (int)(((int64_t)(32 * w + lc) * lda + lr) * 4);
  auto load_fast = [&](float (&x)[32], int64_t blk, int kc) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void*)(A + blk * 256 * lda), 0, (int)(256 * lda * 4), 0x00020000);
    const int soff = ((kc + kofs) % nk) * 256;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const f32x4 v = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, aoff, soff + (int)(16 * u * lda), 0));
      x[4 * u] = v[0]; x[4 * u + 1] = v[1]; x[4 * u + 2] = v[2]; x[4 * u + 3] = v[3];
    }
  };
  auto load_guarded = [&](float (&x)[32], int64_t blk, int kc) {
    const int i = ((kc + kofs) % nk) * 64 + lr;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t jj = blk * 256 + 32 * w + 4 * u + lc;
      if (jj < n) {
        const float* p = A + jj * lda + i;
        if (vec && i + 3 < m) {
          const f32x4 v = *reinterpret_cast<const f32x4*>(p);
          x[4 * u] = v[0]; x[4 * u + 1] = v[1]; x[4 * u + 2] = v[2]; x[4 * u + 3] = v[3];
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) x[4 * u + e] = (i + e < m) ? p[e] : 0.0f;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) x[4 * u + e] = 0.0f;
      }
    }
  };
  float* sw = stg[w];
  auto stage = [&](const float (&x)[32]) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      f32x4 v;
      v[0] = x[4 * u]; v[1] = x[4 * u + 1]; v[2] = x[4 * u + 2]; v[3] = x[4 * u + 3];
      *reinterpret_cast<f32x4*>(sw + (4 * u + lc) * SP + lr) = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  auto compute = [&](f32x16 (&acc)[LT], const float (&xin)[32], const float* q) {
    stage(xin);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      // B operand: rows 32 b + 8 t + [0,8) of column j, from the staged tile
      float xb[8];
      {
        const float* sp = sw + jl * SP + 32 * b + 8 * t;
        const f32x4 u0 = *reinterpret_cast<const f32x4*>(sp);
        const f32x4 u1 = *reinterpret_cast<const f32x4*>(sp + 4);
        xb[0] = u0[0]; xb[1] = u0[1]; xb[2] = u0[2]; xb[3] = u0[3];
        xb[4] = u1[0]; xb[5] = u1[1]; xb[6] = u1[2]; xb[7] = u1[3];
      }
      if constexpr (AR == AR_FP32) {
#pragma unroll
        for (int lt = 0; lt < LT; ++lt) {
          const float* qp = q + (32 * lt + jl) * QP + 32 * b + 8 * t;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            // MFMA k index = lane half b; K-step (t, u) pairs row 32 b + 8 t + u of both operands
            acc[lt] = mfma_f32(qp[u], xb[u], acc[lt]);
          }
        }
      } else if constexpr (AR == AR_FP16) {
        f16x8 bh;
        float xsc[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xsc[u] = xb[u] * sa;
        round8h(xsc, bh);
#pragma unroll
        for (int lt = 0; lt < LT; ++lt) {
          const float* qp = q + (32 * lt + jl) * QP + 32 * b + 8 * t;
          float qv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) qv[u] = qp[u];
          f16x8 ah;
          round8h(qv, ah);
          acc[lt] = mfma_f16(ah, bh, acc[lt]);
        }
      } else {
        bf16x8 bh, bl;
        if (AR == AR_BF16X3 || AR == AR_BF16X4) split8(xb, bh, bl);
        else round8(xb, bh);
#pragma unroll
        for (int lt = 0; lt < LT; ++lt) {
          const __bf16* qh = reinterpret_cast<const __bf16*>(q);
          const int o = (32 * lt + jl) * QP + 32 * b + 8 * t;
          const bf16x8 ah = *reinterpret_cast<const bf16x8*>(qh + o);
          bf16x8 al;
          if (AR == AR_BF16X3 || AR == AR_BF16X4) al = *reinterpret_cast<const bf16x8*>(qh + KP * QP + o);
          acc[lt] = mfma_bf16(ah, bh, acc[lt]);
          if (AR == AR_BF16X3 || AR == AR_BF16X4) {
            acc[lt] = mfma_bf16(al, bh, acc[lt]);
            acc[lt] = mfma_bf16(ah, bl, acc[lt]);
          }
          if (AR == AR_BF16X4) acc[lt] = mfma_bf16(al, bl, acc[lt]);
        }
      }
    }
  };

  // workgroup barrier for the Q chunks in LDS only: their writes complete (lgkmcnt(0)), then
  // s_barrier.  __syncthreads would also wait for the next A chunk's loads (vmcnt(0)) and drain
  // the wave's only chunk in flight at every Q chunk.
  auto q_barrier = [] {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0); vmcnt, expcnt unchanged
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };

  // a finished block: FP16 unscale, the outputs, the column maxima and the Gram of its D tile
  auto epilogue = [&](int64_t blk, f32x16 (&acc)[LT]) {
    const int64_t j = blk * 256 + 32 * w + jl;
    if (AR == AR_FP16 && scales) {
      const float inv = 1.0f / sa;
#pragma unroll
      for (int lt = 0; lt < LT; ++lt)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[lt][v] *= inv;
    }
    if (out_mode == 2) {
      // TR first-step remainder in the 231 layout (trdec6_singleGPU.cu:49-58, 865-877): row c = a + r0 b
      // (b < r1c) of column j goes to out[b + r1c (j + n a)] (r0 = ldo, r1c = r_out).  For each a the
      // wave's 32 columns fill r1c * 32 consecutive floats: the tile goes through the wave's staging
      // area and every a is written as one contiguous run
      const int r0 = (int)ldo, r1c = r_out;
      constexpr int DP = 33;
      static_assert(KP * DP <= 32 * SP, "the D tile fits the staging area");
      float* d = stg[w];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int lt = 0; lt < LT; ++lt)
#pragma unroll
        for (int v = 0; v < 16; ++v) d[(32 * lt + mfma32_row(v, lane)) * DP + jl] = acc[lt][v];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int64_t j0 = blk * 256 + 32 * w;
      const int ncols = (n - j0 >= 32) ? 32 : (n > j0 ? (int)(n - j0) : 0);
      for (int a2 = 0; a2 < r0; ++a2) {
        float* o = out + (int64_t)r1c * (j0 + n * a2);
        for (int e = lane; e < r1c * ncols; e += 64) {
          const int b2 = e % r1c, jj = e / r1c;
          o[e] = d[(a2 + r0 * b2) * DP + jj];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
    } else {
      if constexpr (COLMAX) {
        // columns past n hold zeros
#pragma unroll
        for (int lt = 0; lt < LT; ++lt)
#pragma unroll
          for (int v = 0; v < 16; ++v) amax[lt][v] = fmaxf(amax[lt][v], fabsf(acc[lt][v]));
      }
    }
    if (out_mode != 2 && j < n) {
      if (out_mode == 0) {
        // Z = D^T, row-major n x ldo: kp columns, or the first r_out (a multiple of 4)
        float* o = out + j * ldo;
        const int zc = r_out > 0 ? r_out : KP;
#pragma unroll
        for (int lt = 0; lt < LT; ++lt)
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            if (32 * lt + 8 * g + 4 * b >= zc) continue;
            f32x4 v;
            v[0] = acc[lt][4 * g]; v[1] = acc[lt][4 * g + 1]; v[2] = acc[lt][4 * g + 2]; v[3] = acc[lt][4 * g + 3];
            *reinterpret_cast<f32x4*>(o + 32 * lt + 8 * g + 4 * b) = v;
          }
      } else {
        // next unfolding, column-major r_out x n
        float* o = out + j * (int64_t)r_out;
#pragma unroll
        for (int lt = 0; lt < LT; ++lt)
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int c = 32 * lt + mfma32_row(v, lane);
            if (c < r_out) o[c] = acc[lt][v];
          }
      }
    }
    if constexpr (GRAM) {
      // tile D (kp x 32 columns, zero for columns j >= n) -> the wave's own LDS dt[j][c] (a wave-level
      // fence suffices: no other wave reads it), then G += D D^T with MFMA 32x32x2: A = D[c][j] (lane:
      // c = l%32, j = 2 kk + l/32), B = the same.  Row pitch KP + 1: the 32 lanes of a column write to
      // distinct banks
      constexpr int DP = KP + 1;
      static_assert(32 * DP <= 32 * SP, "D tile fits the staging area");
      float* d = stg[w];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int lt = 0; lt < LT; ++lt)
#pragma unroll
        for (int v = 0; v < 16; ++v) d[jl * DP + 32 * lt + mfma32_row(v, lane)] = acc[lt][v];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) {
        const int jr = 2 * kk + b;
        float x[LT];
#pragma unroll
        for (int lt = 0; lt < LT; ++lt) x[lt] = d[jr * DP + 32 * lt + jl];
        int pidx = 0;
#pragma unroll
        for (int ta = 0; ta < LT; ++ta)
#pragma unroll
          for (int tb = ta; tb < LT; ++tb) { gacc[pidx] = mfma_f32(x[ta], x[tb], gacc[pidx]); ++pidx; }
      }
    }
#pragma unroll
    for (int lt = 0; lt < LT; ++lt) acc[lt] = zero16();
  };

  // Interior blocks (this workgroup's blocks B.x + t gx, t < cnt, all below nfast) run as one chunk
  // sequence g = t nk + kc, software-pipelined across block boundaries: the next chunk's A and Q loads
  // are issued unconditionally before this chunk's MFMAs (the index is clamped at the end of the
  // sequence, where the redundant reload of the last chunk comes from L2).  A load issued under a
  // branch leaves the compiler unable to count the loads in flight at the merge, and it then waits
  // for all of them (vmcnt(0)) before the current chunk, which serialised load and compute.
  // (the buffer offsets are 32-bit: a block of A and the whole of Q stay below 2^31 bytes)
  const bool offs32 = 256 * lda * 4 < (int64_t(1) << 31) && (int64_t)KP * ldq * 4 < (int64_t(1) << 31);
  const int64_t nfast = (rows_whole && offs32) ? n / 256 : 0;
  const int64_t cnt = (B.x < nfast) ? (nfast - 1 - B.x) / B.gx + 1 : 0;
  if (cnt > 0) {
    const int64_t total = cnt * nk;
    using whole_t = std::true_type;
    f32x16 acc[LT];
#pragma unroll
    for (int lt = 0; lt < LT; ++lt) acc[lt] = zero16();
    float x0[32], x1[32];
    f32x4 qr[QF];
    // chunk g: block B.x + (g / nk) gx, K chunk g % nk (tracked incrementally)
    int kc0 = 0;
    int64_t blk0 = B.x;
    auto succ = [&](int kc, int64_t bk, int& kn, int64_t& bn) {
      kn = (kc + 1 == nk) ? 0 : kc + 1;
      bn = (kc + 1 == nk) ? bk + B.gx : bk;
    };
    qload(qr, 0, whole_t{});
    load_fast(x0, blk0, 0);
    qstore(qr, qs[0]);
    q_barrier();
    for (int64_t g = 0; g < total; g += 2) {
      int kc1;
      int64_t blk1;
      succ(kc0, blk0, kc1, blk1);
      if (g + 1 >= total) { kc1 = kc0; blk1 = blk0; }
      qload(qr, kc1, whole_t{});
      load_fast(x1, blk1, kc1);
      compute(acc, x0, qs[0]);
      __builtin_amdgcn_sched_barrier(0);
      qstore(qr, qs[1]);
      q_barrier();
      if (kc0 == nk - 1) epilogue(blk0, acc);
      if (g + 1 >= total) break;
      int kc2;
      int64_t blk2;
      succ(kc1, blk1, kc2, blk2);
      if (g + 2 >= total) { kc2 = kc1; blk2 = blk1; }
      qload(qr, kc2, whole_t{});
      load_fast(x0, blk2, kc2);
      compute(acc, x1, qs[1]);
      __builtin_amdgcn_sched_barrier(0);
      qstore(qr, qs[0]);
      q_barrier();
      if (kc1 == nk - 1) epilogue(blk1, acc);
      kc0 = kc2;
      blk0 = blk2;
    }
  }
  // the other blocks (the tail block, or every block when the rows are not whole): guarded loads
  for (int64_t blk = B.x + cnt * B.gx; blk < blocks_total; blk += B.gx) {
    using guarded_t = std::false_type;
    f32x16 acc[LT];
#pragma unroll
    for (int lt = 0; lt < LT; ++lt) acc[lt] = zero16();
    float x0[32], x1[32];
    f32x4 qr[QF];
    qload(qr, 0, guarded_t{});
    load_guarded(x0, blk, 0);
    qstore(qr, qs[0]);
    q_barrier();
    for (int kc = 0; kc < nk; kc += 2) {
      if (kc + 1 < nk) {
        qload(qr, kc + 1, guarded_t{});
        load_guarded(x1, blk, kc + 1);
      }
      compute(acc, x0, qs[0]);
      __builtin_amdgcn_sched_barrier(0);
      if (kc + 1 < nk) qstore(qr, qs[1]);
      q_barrier();
      if (kc + 1 >= nk) break;
      if (kc + 2 < nk) {
        qload(qr, kc + 2, guarded_t{});
        load_guarded(x0, blk, kc + 2);
      }
      compute(acc, x1, qs[1]);
      __builtin_amdgcn_sched_barrier(0);
      if (kc + 2 < nk) qstore(qr, qs[0]);
      q_barrier();
    }
    epilogue(blk, acc);
  }
  if constexpr (COLMAX) {
    // the lanes of a half share their rows c: reduce over the 32 columns, then over the waves in LDS
    __syncthreads();
    const int jl0 = threadIdx.x & 31;
#pragma unroll
    for (int lt = 0; lt < LT; ++lt)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        float x = amax[lt][v];
#pragma unroll
        for (int off = 1; off < 32; off <<= 1) x = fmaxf(x, __shfl_xor(x, off));
        if (jl0 == 0) atomicMax(&cmx[32 * lt + mfma32_row(v, threadIdx.x & 63)], __float_as_uint(x));
      }
    __syncthreads();
    if (threadIdx.x < KP) atomicMax(&colmax[threadIdx.x], cmx[threadIdx.x]);
  }
  if constexpr (GRAM) {
    for (int e = threadIdx.x; e < KP * KP; e += 512) gs[e] = 0.0;
    __syncthreads();
    for (int ww = 0; ww < 8; ++ww) {
      if (w == ww) {
        int pidx = 0;
#pragma unroll
        for (int ta = 0; ta < LT; ++ta)
#pragma unroll
          for (int tb = ta; tb < LT; ++tb) {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
              const int c = 32 * ta + mfma32_row(v, lane), c2 = 32 * tb + jl;
              gs[c * KP + c2] += (double)gacc[pidx][v];
              if (ta != tb) gs[c2 * KP + c] += (double)gacc[pidx][v];
            }
            ++pidx;
          }
      }
      __syncthreads();
    }
    double* o = gpart + (int64_t)B.x * kg * kg;
    for (int e = threadIdx.x; e < kg * kg; e += 512) o[e] = gs[(e / kg) * KP + (e % kg)];
  }
}

template <int LT, int AR, bool GRAM, bool COLMAX = false>
__global__ __launch_bounds__(512) void k_proj(const float* __restrict__ A, int64_t lda, int m, int64_t n,
                                              const float* __restrict__ Qc, int64_t ldq,
                                              int out_mode, float* __restrict__ out, int64_t ldo, int r_out,
                                              int64_t blocks_total, double* __restrict__ gpart, int kg,
                                              const float* __restrict__ scales, unsigned* __restrict__ colmax) {
  proj_body<LT, AR, GRAM, COLMAX>(A, lda, m, n, Qc, ldq, out_mode, out, ldo, r_out, blocks_total, gpart, kg, scales,
                                  hw_blk(), colmax);
}
template <int LT, int AR, bool GRAM>
__global__ __launch_bounds__(512) void k_proj_grp(const JobProj* __restrict__ jobs, const int2* __restrict__ map) {
  Blk B;
  const JobProj& j = grp_job(jobs, map, B);
  proj_body<LT, AR, GRAM>(j.A, j.lda, j.m, j.n, j.Qc, j.ldq, j.mode, j.out, j.ldo, j.r_out, j.blocks, j.gpart, j.kg,
                          nullptr, B, nullptr);
}

// ------------------------------------------------------------------------------------------------
// projection of a short unfolding: D = Q^T A for m <= 64 rows (the exact step's G^T A)
// ------------------------------------------------------------------------------------------------
// k_proj stages 64-row chunks of 256 columns; at m <= 32 half of each chunk is padding and a block's
// loads are a few hundred bytes per column.  Here a wave computes 32 columns at a time with the rows as
// the MFMA's K dimension: lane (n = lane % 32, hb = lane / 32) holds rows 16 kk + 8 hb + [0, 8) of
// column j0 + n (32 contiguous bytes; the tile's 32 columns are one contiguous block of 32 m floats, so
// its loads use every byte of the lines they touch) as the B operand, and the matching rows of Q's
// columns c = 32 lt + n, loaded and split once per wave, as the A operand.  The next tile's loads are
// issued before this tile's MFMAs.  NK = ceil(m / 16) K-steps; rows i >= m are zero.  Output as
// k_proj's mode 1 (r_out x n column-major) or mode 2 (the TR first step's 231 layout).
// Arithmetic: AR_FP32 (v_mfma_f32_32x32x2_f32, eight per K-step) or AR_BF16X4 (both operands split in
// two bf16 terms, four products: hi hi, lo hi, hi lo, lo lo).
template <int LT, int NK, int AR, int MODE>
__device__ __forceinline__ void proj_short_body(const float* __restrict__ A, int64_t lda, int m, int64_t n,
                                                const float* __restrict__ Qc, int64_t ldq,
                                                float* __restrict__ out, int64_t ldo, int r_out, const Blk B) {
  static_assert(AR == AR_FP32 || AR == AR_BF16X4, "short projection: FP32 or BF16X4");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int cn = lane & 31, hb = lane >> 5;
  // Q operand: Q[i][c] for i = 16 kk + 8 hb + t, c = 32 lt + cn (Qc column-major, ldq; kp = 32 LT columns)
  float qv[LT][NK][8];
#pragma unroll
  for (int lt = 0; lt < LT; ++lt)
#pragma unroll
    for (int kk = 0; kk < NK; ++kk)
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int i = 16 * kk + 8 * hb + t;
        qv[lt][kk][t] = (i < m) ? Qc[(int64_t)(32 * lt + cn) * ldq + i] : 0.0f;
      }
  [[maybe_unused]] bf16x8 qh[LT][NK], ql[LT][NK];
  if constexpr (AR == AR_BF16X4) {
#pragma unroll
    for (int lt = 0; lt < LT; ++lt)
#pragma unroll
      for (int kk = 0; kk < NK; ++kk) split8(qv[lt][kk], qh[lt][kk], ql[lt][kk]);
  }
  const int64_t tiles = (n + 31) / 32;
  // whole 8-row groups as two 16-byte loads: m a multiple of 8 and 16-byte aligned columns
  const bool vec = ((m & 7) == 0) && ((lda & 3) == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0);
  auto load = [&](float (&x)[NK][8], int64_t tile) {
    const int64_t j = tile * 32 + cn;
    const float* col = A + j * lda;
#pragma unroll
    for (int kk = 0; kk < NK; ++kk) {
      const int i0 = 16 * kk + 8 * hb;
      if (vec) {
        if (j < n && i0 < m) {
          const f32x4 u0 = *reinterpret_cast<const f32x4*>(col + i0);
          const f32x4 u1 = *reinterpret_cast<const f32x4*>(col + i0 + 4);
          x[kk][0] = u0[0]; x[kk][1] = u0[1]; x[kk][2] = u0[2]; x[kk][3] = u0[3];
          x[kk][4] = u1[0]; x[kk][5] = u1[1]; x[kk][6] = u1[2]; x[kk][7] = u1[3];
        } else {
#pragma unroll
          for (int t = 0; t < 8; ++t) x[kk][t] = 0.0f;
        }
      } else {
#pragma unroll
        for (int t = 0; t < 8; ++t) x[kk][t] = (j < n && i0 + t < m) ? col[i0 + t] : 0.0f;
      }
    }
  };
  auto compute = [&](const float (&x)[NK][8], f32x16 (&acc)[LT]) {
#pragma unroll
    for (int kk = 0; kk < NK; ++kk) {
      if constexpr (AR == AR_FP32) {
#pragma unroll
        for (int lt = 0; lt < LT; ++lt)
#pragma unroll
          for (int t = 0; t < 8; ++t) acc[lt] = mfma_f32(qv[lt][kk][t], x[kk][t], acc[lt]);
      } else {
        bf16x8 xh, xl;
        split8(x[kk], xh, xl);
#pragma unroll
        for (int lt = 0; lt < LT; ++lt) {
          acc[lt] = mfma_bf16(qh[lt][kk], xh, acc[lt]);
          acc[lt] = mfma_bf16(ql[lt][kk], xh, acc[lt]);
          acc[lt] = mfma_bf16(qh[lt][kk], xl, acc[lt]);
          acc[lt] = mfma_bf16(ql[lt][kk], xl, acc[lt]);
        }
      }
    }
  };
  const bool vout = (MODE == 1) && ((r_out & 3) == 0) && ((reinterpret_cast<uintptr_t>(out) & 15) == 0);
  // mode 2: the tile goes through the wave's LDS area (32 LT rows x 32 columns, pitch 33) so that every
  // a writes its r1c x 32 floats as one contiguous run
  constexpr int DP = 33;
  __shared__ float dstage[MODE == 2 ? 4 * 32 * LT * DP : 1];
  auto store = [&](const f32x16 (&acc)[LT], int64_t tile) {
    if constexpr (MODE == 1) {
      const int64_t j = tile * 32 + cn;
      if (j >= n) return;
      float* o = out + j * (int64_t)r_out;
#pragma unroll
      for (int lt = 0; lt < LT; ++lt)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int c0 = 32 * lt + 8 * g + 4 * hb;   // rows c0 .. c0 + 3 of D in acc[lt][4 g .. 4 g + 3]
          if (vout && c0 + 3 < r_out) {
            *reinterpret_cast<f32x4*>(o + c0) =
                f32x4{acc[lt][4 * g], acc[lt][4 * g + 1], acc[lt][4 * g + 2], acc[lt][4 * g + 3]};
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (c0 + e < r_out) o[c0 + e] = acc[lt][4 * g + e];
          }
        }
    } else {
      // row c = a + r0 b (b < r1c) of column j to out[b + r1c (j + n a)] (trdec6_singleGPU.cu:49-58)
      float* d = dstage + w * (32 * LT * DP);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int lt = 0; lt < LT; ++lt)
#pragma unroll
        for (int v = 0; v < 16; ++v) d[(32 * lt + mfma32_row(v, lane)) * DP + cn] = acc[lt][v];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int r0 = (int)ldo, r1c = r_out;
      const int64_t j0 = tile * 32;
      const int ncols = (n - j0 >= 32) ? 32 : (int)(n - j0);
      for (int a2 = 0; a2 < r0; ++a2) {
        float* o = out + (int64_t)r1c * (j0 + n * a2);
        for (int e = lane; e < r1c * ncols; e += 64) {
          const int b2 = e % r1c, jj = e / r1c;
          o[e] = d[(a2 + r0 * b2) * DP + jj];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  };
  int64_t tile = (int64_t)B.x * nw + w;
  const int64_t stride = (int64_t)B.gx * nw;
  float x0[NK][8], x1[NK][8];
  if (tile < tiles) load(x0, tile);
  for (; tile < tiles; tile += 2 * stride) {
    // even tile from x0, the next one's loads in flight; then the odd tile from x1
    if (tile + stride < tiles) load(x1, tile + stride);
    f32x16 acc[LT];
#pragma unroll
    for (int lt = 0; lt < LT; ++lt) acc[lt] = zero16();
    compute(x0, acc);
    store(acc, tile);
    if (tile + stride >= tiles) break;
    if (tile + 2 * stride < tiles) load(x0, tile + 2 * stride);
#pragma unroll
    for (int lt = 0; lt < LT; ++lt) acc[lt] = zero16();
    compute(x1, acc);
    store(acc, tile + stride);
  }
}
template <int LT, int NK, int AR, int MODE>
__global__ __launch_bounds__(256) void k_proj_short(const float* __restrict__ A, int64_t lda, int m, int64_t n,
                                                    const float* __restrict__ Qc, int64_t ldq,
                                                    float* __restrict__ out, int64_t ldo, int r_out) {
  proj_short_body<LT, NK, AR, MODE>(A, lda, m, n, Qc, ldq, out, ldo, r_out, hw_blk());
}
// grouped: mode 1 jobs (a batch's exact steps)
template <int LT, int NK, int AR>
__global__ __launch_bounds__(256) void k_proj_short_grp(const JobProj* __restrict__ jobs, const int2* __restrict__ map) {
  Blk B;
  const JobProj& j = grp_job(jobs, map, B);
  proj_short_body<LT, NK, AR, 1>(j.A, j.lda, j.m, j.n, j.Qc, j.ldq, j.out, j.ldo, j.r_out, B);
}

// ------------------------------------------------------------------------------------------------
// exact-path Gram: C = A A^T for m <= 64 (MT = number of 32-row tiles), FP64 partial per WG
// ------------------------------------------------------------------------------------------------
template <int MT, int AR>
__device__ __forceinline__ void gram_small_body(const float* __restrict__ A, int64_t lda, int m, int64_t n,
                                                int64_t cols_per_wg, double* __restrict__ part,
                                                float* __restrict__ absmax, const Blk B) {
  constexpr int NP = MT * (MT + 1) / 2;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int r = lane & 31;
  const int b = lane >> 5;
  const int64_t j0 = (int64_t)B.x * cols_per_wg;
  const int64_t j1 = (j0 + cols_per_wg < n) ? j0 + cols_per_wg : n;

  double dacc[NP][16];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int v = 0; v < 16; ++v) dacc[p][v] = 0.0;
  float amax = 0.0f;

  // each wave takes 16-column steps j0 + 16 (w + 4 i); flush the fp32 MFMA accumulator to fp64
  // every 64 steps
  int64_t step = 0;
  f32x16 acc[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) acc[p] = zero16();
  // the loads of step jc + 64 are issued before the MFMAs of step jc (register double buffer)
  auto loadv = [&](float (&v)[MT][8], int64_t jc) {
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int64_t j = jc + 8 * b + t;
        const int i = 32 * mt + r;
        v[mt][t] = (j < j1 && i < m) ? A[i + j * lda] : 0.0f;
      }
  };
  float vn[MT][8];
  loadv(vn, j0 + 16 * w);
  for (int64_t jc = j0 + 16 * w; jc < j1; jc += 64) {
    float v[MT][8];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        v[mt][t] = vn[mt][t];
        amax = fmaxf(amax, fabsf(v[mt][t]));
      }
    if (jc + 64 < j1) loadv(vn, jc + 64);
    if constexpr (AR == AR_FP32) {
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        int p = 0;
#pragma unroll
        for (int m1 = 0; m1 < MT; ++m1)
#pragma unroll
          for (int m2 = m1; m2 < MT; ++m2) { acc[p] = mfma_f32(v[m1][t], v[m2][t], acc[p]); ++p; }
      }
    } else {
      bf16x8 hi[MT], lo[MT];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        if (AR == AR_BF16X3 || AR == AR_BF16X4) split8(v[mt], hi[mt], lo[mt]);
        else round8(v[mt], hi[mt]);
      }
      int p = 0;
#pragma unroll
      for (int m1 = 0; m1 < MT; ++m1)
#pragma unroll
        for (int m2 = m1; m2 < MT; ++m2) {
          acc[p] = mfma_bf16(hi[m1], hi[m2], acc[p]);
          if (AR == AR_BF16X3 || AR == AR_BF16X4) {
            acc[p] = mfma_bf16(lo[m1], hi[m2], acc[p]);
            acc[p] = mfma_bf16(hi[m1], lo[m2], acc[p]);
          }
          if (AR == AR_BF16X4) acc[p] = mfma_bf16(lo[m1], lo[m2], acc[p]);
          ++p;
        }
    }
    if (++step == 64) {
      step = 0;
#pragma unroll
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int q = 0; q < 16; ++q) dacc[p][q] += (double)acc[p][q];
        acc[p] = zero16();
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int q = 0; q < 16; ++q) dacc[p][q] += (double)acc[p][q];

  // combine the 4 waves through LDS; one slab per WG: part[block][m][m]
  __shared__ double red[4][NP][1024];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int v = 0; v < 16; ++v) red[w][p][mfma32_row(v, lane) * 32 + r] = dacc[p][v];
  __syncthreads();
  double* o = part + (int64_t)B.x * m * m;
  for (int e = threadIdx.x; e < NP * 1024; e += 256) {
    const int p = e >> 10, rc = e & 1023;
    const double sum = red[0][p][rc] + red[1][p][rc] + red[2][p][rc] + red[3][p][rc];
    int m1 = 0, m2 = 0;
    if (MT == 2) { m1 = (p == 2) ? 1 : 0; m2 = (p == 0) ? 0 : 1; }
    const int i = 32 * m1 + (rc >> 5), k = 32 * m2 + (rc & 31);
    if (i < m && k < m) {
      o[i * m + k] = sum;
      if (m1 != m2) o[k * m + i] = sum;
    }
  }
  // |A|max for the optional fp16 scaling
  if (absmax) {
    for (int off = 32; off > 0; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off));
    if (lane == 0) atomicMax(reinterpret_cast<unsigned int*>(absmax), __float_as_uint(amax));
  }
}

template <int MT, int AR>
__global__ __launch_bounds__(256) void k_gram_small(const float* __restrict__ A, int64_t lda, int m, int64_t n,
                                                    int64_t cols_per_wg, double* __restrict__ part,
                                                    float* __restrict__ absmax) {
  gram_small_body<MT, AR>(A, lda, m, n, cols_per_wg, part, absmax, hw_blk());
}
template <int MT, int AR>
__global__ __launch_bounds__(256) void k_gram_small_grp(const JobGramSmall* __restrict__ jobs,
                                                        const int2* __restrict__ map) {
  Blk B;
  const JobGramSmall& j = grp_job(jobs, map, B);
  gram_small_body<MT, AR>(j.A, j.lda, j.m, j.n, j.cpw, j.part, nullptr, B);
}

// ------------------------------------------------------------------------------------------------
// deterministic slab reductions
// ------------------------------------------------------------------------------------------------
// out[e] = sum_s part[s * count + e] in a fixed order: WG handles 32 elements; its 8 thread groups
// take slabs g, g + 8, ...; the 8 group sums are added in group order.
template <typename TI, typename TO>
__device__ __forceinline__ void reduce_body(const TI* __restrict__ part, int64_t S, int64_t count,
                                            TO* __restrict__ out, const Blk B) {
  __shared__ double red[8][33];
  const int el = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int64_t e = (int64_t)B.x * 32 + el;
  double acc = 0.0;
  // unrolled: sixteen independent loads in flight per thread, added in slab order
  if (e < count) {
#pragma unroll 16
    for (int64_t k = g; k < S; k += 8) acc += (double)part[k * count + e];
  }
  red[g][el] = acc;
  __syncthreads();
  if (g == 0 && e < count) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) t += red[i][el];
    out[e] = (TO)t;
  }
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_reduce(const TI* __restrict__ part, int64_t S, int64_t count,
                                                TO* __restrict__ out) {
  reduce_body<TI, TO>(part, S, count, out, hw_blk());
}
template <typename T>
__global__ __launch_bounds__(256) void k_reduce_grp(const JobReduce* __restrict__ jobs, const int2* __restrict__ map) {
  Blk B;
  const JobReduce& j = grp_job(jobs, map, B);
  reduce_body<T, T>(static_cast<const T*>(j.part), j.S, j.count, static_cast<T*>(j.out), B);
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
static int g_num_cus = 0;
// direct-to-LDS sketch (k_sketch_dl); TTX_SKETCH_DL=0 selects the register-buffered kernel
static const bool g_sketch_dl = [] {
  const char* e = getenv("TTX_SKETCH_DL");
  return !(e && e[0] == '0');
}();
int num_cus() {
  if (g_num_cus == 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) g_num_cus = prop.multiProcessorCount;
    if (g_num_cus <= 0) g_num_cus = 256;
  }
  return g_num_cus;
}

SketchPlan plan_sketch(int m, int64_t n, int kp, int wgs_hint) {
  SketchPlan p;
  const int NW = 8;                      // waves per WG
  const int RW = (kp == 32) ? 128 : 64;  // rows per wave (32 x 4 / LT)
  const int rows = m < RW * NW ? m : RW * NW;
  int rb = (rows + RW - 1) / RW;
  int RB = 1;
  while (RB < rb) RB *= 2;
  p.RB = RB;
  p.ysplit = (m + RW * RB - 1) / (RW * RB);
  int CG = NW / RB;
  const int cg_lds = (kp == 32) ? 8 : 4;   // keep the two X chunks within 64 KB of LDS
  if (CG > cg_lds) CG = cg_lds;
  p.CG = CG;
  const int KC = 16 * CG;
  const int64_t nchunks = (n + KC - 1) / KC;
  const int waves = RB * CG;
  // two full-width WGs per CU; narrower WGs share a CU
  int64_t target = wgs_hint > 0 ? wgs_hint : (int64_t)num_cus() * 2 * (NW / waves) / p.ysplit;
  if (target < 1) target = 1;
  int64_t cpw = (nchunks + target - 1) / target;
  if (cpw < 1) cpw = 1;
  p.chunks_per_wg = cpw;
  p.grid_x = (int)((nchunks + cpw - 1) / cpw);
  if (p.grid_x < 1) p.grid_x = 1;
  p.slabs = (int64_t)p.grid_x * CG;
  p.kp = kp;
  p.lds_bytes = 2 * kp * (KC + 4) * (int)sizeof(float);
  return p;
}

SketchPlan plan_sketch_batch(int m, int64_t n, int kp, int64_t cpw) {
  SketchPlan p = plan_sketch(m, n, kp, 1);
  const int64_t nchunks = (n + 16 * p.CG - 1) / (16 * p.CG);
  p.chunks_per_wg = cpw < 1 ? 1 : cpw;
  p.grid_x = (int)((nchunks + p.chunks_per_wg - 1) / p.chunks_per_wg);
  if (p.grid_x < 1) p.grid_x = 1;
  p.slabs = (int64_t)p.grid_x * p.CG;
  p.plain = true;
  return p;
}

template <int LT, int AR, bool OM>
static void launch_sketch_t(const SketchPlan& p, const float* A, int64_t lda, int m, int64_t n, const float* X,
                            int64_t ldx, uint64_t seed, uint32_t stream, int ell, float* part, hipStream_t st,
                            const float* scales) {
  dim3 grid(p.grid_x, p.ysplit);
  const int fq = (32 * LT) / (16 * p.RB);   // float4 pieces of the X chunk per thread
  if constexpr (OM) {
    // A through LDS, as many chunks deep (4 at most) as fit, when the rows are 16-byte aligned
    constexpr int RW = 128 / LT;
    const int nw = p.RB * p.CG;
    const int tile_bytes = 16 * RW * (int)sizeof(float) * nw;
    constexpr int LDS_MAX = 160 * 1024;
    if ((lda & 3) == 0 && (m & 3) == 0 && g_sketch_dl && !p.plain) {
      if (p.lds_bytes + 4 * tile_bytes <= LDS_MAX) {
        hipLaunchKernelGGL((k_sketch_dl<LT, AR, 4, false>), grid, dim3(64 * nw), p.lds_bytes + 4 * tile_bytes, st, A, lda, m,
                           n, seed, stream, ell, p.RB, p.CG, p.chunks_per_wg, part, scales, nullptr);
        return;
      }
      if (p.lds_bytes + 3 * tile_bytes <= LDS_MAX) {
        hipLaunchKernelGGL((k_sketch_dl<LT, AR, 3, false>), grid, dim3(64 * nw), p.lds_bytes + 3 * tile_bytes, st, A, lda, m,
                           n, seed, stream, ell, p.RB, p.CG, p.chunks_per_wg, part, scales, nullptr);
        return;
      }
      if (p.lds_bytes + 2 * tile_bytes <= LDS_MAX) {
        hipLaunchKernelGGL((k_sketch_dl<LT, AR, 2, false>), grid, dim3(64 * nw), p.lds_bytes + 2 * tile_bytes, st, A, lda, m,
                           n, seed, stream, ell, p.RB, p.CG, p.chunks_per_wg, part, scales, nullptr);
        return;
      }
    }
    hipLaunchKernelGGL((k_sketch<LT, AR, OM, 1>), grid, dim3(64 * p.RB * p.CG), p.lds_bytes, st, A, lda, m, n, X, ldx,
                       seed, stream, ell, p.RB, p.CG, p.chunks_per_wg, part, scales);
  }
  else if (fq <= 1)
    hipLaunchKernelGGL((k_sketch<LT, AR, OM, 1>), grid, dim3(64 * p.RB * p.CG), p.lds_bytes, st, A, lda, m, n, X, ldx,
                       seed, stream, ell, p.RB, p.CG, p.chunks_per_wg, part, scales);
  else if (fq == 2)
    hipLaunchKernelGGL((k_sketch<LT, AR, OM, 2>), grid, dim3(64 * p.RB * p.CG), p.lds_bytes, st, A, lda, m, n, X, ldx,
                       seed, stream, ell, p.RB, p.CG, p.chunks_per_wg, part, scales);
  else
    hipLaunchKernelGGL((k_sketch<LT, AR, OM, 4>), grid, dim3(64 * p.RB * p.CG), p.lds_bytes, st, A, lda, m, n, X, ldx,
                       seed, stream, ell, p.RB, p.CG, p.chunks_per_wg, part, scales);
}

template <int LT, bool OM>
static void launch_sketch_ar(int ar, const SketchPlan& p, const float* A, int64_t lda, int m, int64_t n, const float* X,
                             int64_t ldx, uint64_t seed, uint32_t stream, int ell, float* part, hipStream_t st,
                             const float* sc) {
  switch (ar) {
    case AR_FP32: launch_sketch_t<LT, AR_FP32, OM>(p, A, lda, m, n, X, ldx, seed, stream, ell, part, st, sc); break;
    case AR_BF16: launch_sketch_t<LT, AR_BF16, OM>(p, A, lda, m, n, X, ldx, seed, stream, ell, part, st, sc); break;
    case AR_FP16: launch_sketch_t<LT, AR_FP16, OM>(p, A, lda, m, n, X, ldx, seed, stream, ell, part, st, sc); break;
    default: launch_sketch_t<LT, AR_BF16X3, OM>(p, A, lda, m, n, X, ldx, seed, stream, ell, part, st, sc); break;
  }
}

bool sketch_gram0_ok(const SketchPlan& p, int m, int64_t lda, int kp) {
  if (kp != 64 || (m % 64) != 0 || (lda & 3) != 0 || p.CG != 1 || !g_sketch_dl) return false;
  const int nw = p.RB * p.CG;
  const int tile_bytes = 4 * (256 + 16) * (int)sizeof(float) * nw;
  return p.lds_bytes + 3 * tile_bytes <= 160 * 1024 && p.RB * 64 * p.ysplit >= m;
}

int64_t sketch_gram0_partials(const SketchPlan& p) { return (int64_t)p.grid_x * p.ysplit; }

int64_t omega_chunks_bytes(int64_t n) { return (n + 15) / 16 * 2048; }

void launch_omega_chunks(uint64_t seed, uint32_t stream, int64_t n, int ell, void* out, hipStream_t st) {
  const int64_t nchunks = (n + 15) / 16;
  int64_t grid = (nchunks * 16 + 255) / 256;
  const int64_t cap = (int64_t)num_cus() * 32;
  if (grid > cap) grid = cap;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_omega_chunks, dim3((unsigned)grid), dim3(256), 0, st, seed, stream, n, ell, nchunks,
                     reinterpret_cast<__bf16*>(out));
}

void launch_sketch_gram0(const SketchPlan& p, const float* A, int64_t lda, int m, int64_t n, uint64_t seed,
                         uint32_t stream, int ell, float* part, double* gpart, hipStream_t st, const void* omega,
                         bool gram3) {
  dim3 grid(p.grid_x, p.ysplit);
  const int nw = p.RB * p.CG;
  const int tile_bytes = 4 * (256 + 16) * (int)sizeof(float) * nw;
  const int lds = p.lds_bytes + (p.lds_bytes + 4 * tile_bytes <= 160 * 1024 ? 4 : 3) * tile_bytes;
  const __bf16* om = reinterpret_cast<const __bf16*>(omega);
  if (om && gram3 && p.lds_bytes + 4 * tile_bytes <= 160 * 1024)
    hipLaunchKernelGGL((k_sketch_dl<2, AR_BF16X3, 4, true, true, true>), grid, dim3(64 * nw), lds, st, A, lda, m, n,
                       seed, stream, ell, p.RB, p.CG, p.chunks_per_wg, part, (const float*)nullptr, gpart, om);
  else if (om && gram3)
    hipLaunchKernelGGL((k_sketch_dl<2, AR_BF16X3, 3, true, true, true>), grid, dim3(64 * nw), lds, st, A, lda, m, n,
                       seed, stream, ell, p.RB, p.CG, p.chunks_per_wg, part, (const float*)nullptr, gpart, om);
  else if (om && p.lds_bytes + 4 * tile_bytes <= 160 * 1024)
    hipLaunchKernelGGL((k_sketch_dl<2, AR_BF16X3, 4, true, true>), grid, dim3(64 * nw), lds, st, A, lda, m, n, seed,
                       stream, ell, p.RB, p.CG, p.chunks_per_wg, part, (const float*)nullptr, gpart, om);
  else if (om)
    hipLaunchKernelGGL((k_sketch_dl<2, AR_BF16X3, 3, true, true>), grid, dim3(64 * nw), lds, st, A, lda, m, n, seed,
                       stream, ell, p.RB, p.CG, p.chunks_per_wg, part, (const float*)nullptr, gpart, om);
  else if (p.lds_bytes + 4 * tile_bytes <= 160 * 1024)
    hipLaunchKernelGGL((k_sketch_dl<2, AR_BF16X3, 4, true>), grid, dim3(64 * nw), lds, st, A, lda, m, n, seed, stream,
                       ell, p.RB, p.CG, p.chunks_per_wg, part, (const float*)nullptr, gpart);
  else
    hipLaunchKernelGGL((k_sketch_dl<2, AR_BF16X3, 3, true>), grid, dim3(64 * nw), lds, st, A, lda, m, n, seed, stream,
                       ell, p.RB, p.CG, p.chunks_per_wg, part, (const float*)nullptr, gpart);
}

void launch_sketch(int ar, const SketchPlan& p, const float* A, int64_t lda, int m, int64_t n, const float* X,
                   int64_t ldx, uint64_t seed, uint32_t stream, int ell, float* part, hipStream_t st,
                   const float* scales) {
  const bool om = (X == nullptr);
  if (p.kp == 32) {
    if (om) launch_sketch_ar<1, true>(ar, p, A, lda, m, n, X, ldx, seed, stream, ell, part, st, scales);
    else launch_sketch_ar<1, false>(ar, p, A, lda, m, n, X, ldx, seed, stream, ell, part, st, scales);
  } else {
    if (om) launch_sketch_ar<2, true>(ar, p, A, lda, m, n, X, ldx, seed, stream, ell, part, st, scales);
    else launch_sketch_ar<2, false>(ar, p, A, lda, m, n, X, ldx, seed, stream, ell, part, st, scales);
  }
}

// ------------------------------------------------------------------------------------------------
// FP16 operand scales (powers of two): scales[0] for A from ||A||_F^2 = sum of the previous step's
// kept sigma^2 (RMS entry scaled to [8, 16)); scales[1 + c] for column c of X from max_j |X[j][c]|
// (scaled to [128, 256)).  Zero or non-finite inputs give scale 1.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ float pow2_scale_to(float v, int target_exp) {
  if (!(v > 0.0f) || !isfinite(v)) return 1.0f;
  const int e = (int)((__float_as_uint(v) >> 23) & 0xFF) - 127;     // floor(log2 v) for normal v
  int t = target_exp - e;
  t = t < -100 ? -100 : (t > 100 ? 100 : t);
  return ldexpf(1.0f, t);
}

__global__ void k_fp16_ascale(const float* __restrict__ sig, int r, double numel, float* __restrict__ scales) {
  if (threadIdx.x != 0) return;
  double e2 = 0.0;
  if (sig)
    for (int i = 0; i < r; ++i) e2 += (double)sig[i] * sig[i];
  const float rms = (float)sqrt(e2 / numel);
  scales[0] = sig ? pow2_scale_to(rms, 3) : 1.0f;
}

__global__ void k_colmax(const float* __restrict__ X, int64_t n, int64_t ldx, int kp, unsigned* __restrict__ mx) {
  __shared__ unsigned smax[64];
  if (threadIdx.x < 64) smax[threadIdx.x] = 0u;
  __syncthreads();
  const int c = threadIdx.x % kp;
  const int rstep = blockDim.x / kp;
  unsigned local = 0u;
  for (int64_t j = (int64_t)blockIdx.x * rstep + threadIdx.x / kp; j < n; j += (int64_t)gridDim.x * rstep) {
    const unsigned b = __float_as_uint(fabsf(X[j * ldx + c]));
    local = b > local ? b : local;
  }
  atomicMax(&smax[c], local);
  __syncthreads();
  if (threadIdx.x < kp) atomicMax(&mx[threadIdx.x], smax[threadIdx.x]);
}

__global__ void k_colscale(unsigned* __restrict__ mx, int kp, float* __restrict__ scales) {
  const int c = threadIdx.x;
  if (c < kp) {
    scales[1 + c] = pow2_scale_to(__uint_as_float(mx[c]), 7);
    mx[c] = 0u;   // ready for the next pass
  }
}

void launch_fp16_ascale(const float* sig_prev, int r_prev, double numel, float* scales, hipStream_t st) {
  hipLaunchKernelGGL(k_fp16_ascale, dim3(1), dim3(64), 0, st, sig_prev, r_prev, numel, scales);
}

// mx: kp zeroed words (left zeroed again afterwards)
void launch_fp16_colscale(unsigned* mx, int kp, float* scales, hipStream_t st) {
  hipLaunchKernelGGL(k_colscale, dim3(1), dim3(64), 0, st, mx, kp, scales);
}

void launch_fp16_xscale(const float* X, int64_t n, int64_t ldx, int kp, unsigned* mx, float* scales, hipStream_t st) {
  int64_t grid = (n * kp + 255) / 256;
  if (grid > 1024) grid = 1024;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_colmax, dim3((unsigned)grid), dim3(256), 0, st, X, n, ldx, kp, mx);
  hipLaunchKernelGGL(k_colscale, dim3(1), dim3(64), 0, st, mx, kp, scales);
}

int64_t proj_grid(int64_t n) {
  const int64_t blocks = (n + 255) / 256;
  int64_t grid = (int64_t)num_cus() * 2;
  if (grid > blocks) grid = blocks;
  if (grid < 1) grid = 1;
  return grid;
}

int64_t proj_blocks(int64_t n) { return (n + 255) / 256; }

template <int LT, int AR>
static void launch_proj_t(const float* A, int64_t lda, int m, int64_t n, const float* Qc, int64_t ldq, int mode,
                          float* out, int64_t ldo, int r_out, double* gpart, int kg, hipStream_t st,
                          const float* sc, int64_t grid_in = 0, unsigned* colmax = nullptr) {
  const int64_t blocks = proj_blocks(n);
  const int64_t grid = grid_in > 0 ? grid_in : proj_grid(n);
  if (gpart)
    hipLaunchKernelGGL((k_proj<LT, AR, true>), dim3((unsigned)grid), dim3(512), 0, st, A, lda, m, n, Qc, ldq, mode,
                       out, ldo, r_out, blocks, gpart, kg, sc, (unsigned*)nullptr);
  else if (AR == AR_FP16 && colmax && mode == 0)
    hipLaunchKernelGGL((k_proj<LT, AR, false, AR == AR_FP16>), dim3((unsigned)grid), dim3(512), 0, st, A, lda, m, n, Qc,
                       ldq, mode, out, ldo, r_out, blocks, gpart, kg, sc, colmax);
  else
    hipLaunchKernelGGL((k_proj<LT, AR, false>), dim3((unsigned)grid), dim3(512), 0, st, A, lda, m, n, Qc, ldq, mode,
                       out, ldo, r_out, blocks, gpart, kg, sc, (unsigned*)nullptr);
}

template <int LT>
static void launch_proj_ar(int ar, const float* A, int64_t lda, int m, int64_t n, const float* Qc, int64_t ldq,
                           int mode, float* out, int64_t ldo, int r_out, double* gpart, int kg, hipStream_t st,
                           const float* sc, int64_t grid, unsigned* colmax = nullptr) {
  switch (ar) {
    case AR_FP32: launch_proj_t<LT, AR_FP32>(A, lda, m, n, Qc, ldq, mode, out, ldo, r_out, gpart, kg, st, sc, grid); break;
    case AR_BF16: launch_proj_t<LT, AR_BF16>(A, lda, m, n, Qc, ldq, mode, out, ldo, r_out, gpart, kg, st, sc, grid); break;
    case AR_FP16: launch_proj_t<LT, AR_FP16>(A, lda, m, n, Qc, ldq, mode, out, ldo, r_out, gpart, kg, st, sc, grid,
                                             colmax); break;
    case AR_BF16X4: launch_proj_t<LT, AR_BF16X4>(A, lda, m, n, Qc, ldq, mode, out, ldo, r_out, gpart, kg, st, sc, grid); break;
    default: launch_proj_t<LT, AR_BF16X3>(A, lda, m, n, Qc, ldq, mode, out, ldo, r_out, gpart, kg, st, sc, grid); break;
  }
}

void launch_proj(int ar, int kp, const float* A, int64_t lda, int m, int64_t n, const float* Qc, int64_t ldq,
                 int mode, float* out, int64_t ldo, int r_out, hipStream_t st, double* gpart, int kg,
                 const float* scales) {
  if (kp == 32) launch_proj_ar<1>(ar, A, lda, m, n, Qc, ldq, mode, out, ldo, r_out, gpart, kg, st, scales, 0);
  else launch_proj_ar<2>(ar, A, lda, m, n, Qc, ldq, mode, out, ldo, r_out, gpart, kg, st, scales, 0);
}

void launch_proj_grid(int ar, int kp, const float* A, int64_t lda, int m, int64_t n, const float* Qc, int64_t ldq,
                      int mode, float* out, int64_t ldo, int r_out, int64_t grid, hipStream_t st, double* gpart, int kg,
                      const float* scales, unsigned* colmax) {
  if (kp == 32) launch_proj_ar<1>(ar, A, lda, m, n, Qc, ldq, mode, out, ldo, r_out, gpart, kg, st, scales, grid, colmax);
  else launch_proj_ar<2>(ar, A, lda, m, n, Qc, ldq, mode, out, ldo, r_out, gpart, kg, st, scales, grid, colmax);
}

// one 256-thread block per 4 x 32 columns, at most 16 per CU (plain), or 16 tiles per wave (batch)
int64_t proj_short_grid(int64_t n, bool batch) {
  const int64_t tiles = (n + 31) / 32;
  int64_t g = batch ? (tiles + 63) / 64 : (tiles + 3) / 4;
  if (!batch && g > (int64_t)num_cus() * 16) g = (int64_t)num_cus() * 16;
  return g < 1 ? 1 : g;
}

template <int LT, int AR>
static void launch_proj_short_nk(int nk, dim3 g, hipStream_t st, const float* A, int64_t lda, int m, int64_t n,
                                 const float* Qc, int64_t ldq, int mode, float* out, int64_t ldo, int r_out) {
#define TTX_PS(NK, MODE) hipLaunchKernelGGL((k_proj_short<LT, NK, AR, MODE>), g, dim3(256), 0, st, A, lda, m, n, Qc, ldq, \
                                            out, ldo, r_out)
#define TTX_PSM(MODE) { if (nk <= 1) TTX_PS(1, MODE); else if (nk == 2) TTX_PS(2, MODE); \
                        else if (nk == 3) TTX_PS(3, MODE); else TTX_PS(4, MODE); }
  if (mode == 2) TTX_PSM(2) else TTX_PSM(1)
#undef TTX_PSM
#undef TTX_PS
}

void launch_proj_short(int ar, int kp, const float* A, int64_t lda, int m, int64_t n, const float* Qc, int64_t ldq,
                       int mode, float* out, int64_t ldo, int r_out, int64_t grid, hipStream_t st) {
  const int nk = (m + 15) / 16;
  const dim3 g((unsigned)(grid > 0 ? grid : proj_short_grid(n, false)));
  if (ar == AR_FP32) {
    if (kp == 32) launch_proj_short_nk<1, AR_FP32>(nk, g, st, A, lda, m, n, Qc, ldq, mode, out, ldo, r_out);
    else launch_proj_short_nk<2, AR_FP32>(nk, g, st, A, lda, m, n, Qc, ldq, mode, out, ldo, r_out);
  } else {
    if (kp == 32) launch_proj_short_nk<1, AR_BF16X4>(nk, g, st, A, lda, m, n, Qc, ldq, mode, out, ldo, r_out);
    else launch_proj_short_nk<2, AR_BF16X4>(nk, g, st, A, lda, m, n, Qc, ldq, mode, out, ldo, r_out);
  }
}

int64_t gram_small_grid(int64_t n) {
  // four 4-wave WGs per CU (the kernel's registers allow five)
  int64_t g = (int64_t)num_cus() * 4;
  const int64_t minc = 4096;   // at least this many columns per WG
  if (g * minc > n) g = (n + minc - 1) / minc;
  if (g < 1) g = 1;
  return g;
}

void launch_gram_small(int ar, const float* A, int64_t lda, int m, int64_t n, double* part, int64_t grid,
                       float* absmax, hipStream_t st) {
  int64_t cpw = (n + grid - 1) / grid;
  cpw = ((cpw + 63) / 64) * 64;
  if (m <= 32) {
    switch (ar) {
      case AR_FP32: hipLaunchKernelGGL((k_gram_small<1, AR_FP32>), dim3(grid), dim3(256), 0, st, A, lda, m, n, cpw, part, absmax); break;
      case AR_BF16: hipLaunchKernelGGL((k_gram_small<1, AR_BF16>), dim3(grid), dim3(256), 0, st, A, lda, m, n, cpw, part, absmax); break;
      case AR_BF16X4: hipLaunchKernelGGL((k_gram_small<1, AR_BF16X4>), dim3(grid), dim3(256), 0, st, A, lda, m, n, cpw, part, absmax); break;
      default: hipLaunchKernelGGL((k_gram_small<1, AR_BF16X3>), dim3(grid), dim3(256), 0, st, A, lda, m, n, cpw, part, absmax); break;
    }
  } else {
    switch (ar) {
      case AR_FP32: hipLaunchKernelGGL((k_gram_small<2, AR_FP32>), dim3(grid), dim3(256), 0, st, A, lda, m, n, cpw, part, absmax); break;
      case AR_BF16: hipLaunchKernelGGL((k_gram_small<2, AR_BF16>), dim3(grid), dim3(256), 0, st, A, lda, m, n, cpw, part, absmax); break;
      case AR_BF16X4: hipLaunchKernelGGL((k_gram_small<2, AR_BF16X4>), dim3(grid), dim3(256), 0, st, A, lda, m, n, cpw, part, absmax); break;
      default: hipLaunchKernelGGL((k_gram_small<2, AR_BF16X3>), dim3(grid), dim3(256), 0, st, A, lda, m, n, cpw, part, absmax); break;
    }
  }
}

void launch_reduce_f32(const float* part, int64_t S, int64_t count, float* out, hipStream_t st) {
  const int64_t grid = (count + 31) / 32;
  hipLaunchKernelGGL((k_reduce<float, float>), dim3((unsigned)grid), dim3(256), 0, st, part, S, count, out);
}
void launch_reduce_f64(const double* part, int64_t S, int64_t count, double* out, hipStream_t st) {
  const int64_t grid = (count + 31) / 32;
  hipLaunchKernelGGL((k_reduce<double, double>), dim3((unsigned)grid), dim3(256), 0, st, part, S, count, out);
}

// ------------------------------------------------------------------------------------------------
// grouped launchers
// ------------------------------------------------------------------------------------------------
#define TTX_GRP_ARGS dim3((unsigned)g.blocks), dim3(g.threads), g.lds, st
void grp_gram_small(int ar, int mt, const GroupLaunch& g, hipStream_t st) {
  const JobGramSmall* J = static_cast<const JobGramSmall*>(g.jobs);
#define TTX_G(MT, AR) hipLaunchKernelGGL((k_gram_small_grp<MT, AR>), TTX_GRP_ARGS, J, g.map)
  if (mt == 1) {
    switch (ar) {
      case AR_FP32: TTX_G(1, AR_FP32); break;
      case AR_BF16: TTX_G(1, AR_BF16); break;
      case AR_BF16X4: TTX_G(1, AR_BF16X4); break;
      default: TTX_G(1, AR_BF16X3); break;
    }
  } else {
    switch (ar) {
      case AR_FP32: TTX_G(2, AR_FP32); break;
      case AR_BF16: TTX_G(2, AR_BF16); break;
      case AR_BF16X4: TTX_G(2, AR_BF16X4); break;
      default: TTX_G(2, AR_BF16X3); break;
    }
  }
#undef TTX_G
}

void grp_reduce(bool f64, const GroupLaunch& g, hipStream_t st) {
  const JobReduce* J = static_cast<const JobReduce*>(g.jobs);
  if (f64) hipLaunchKernelGGL(k_reduce_grp<double>, TTX_GRP_ARGS, J, g.map);
  else hipLaunchKernelGGL(k_reduce_grp<float>, TTX_GRP_ARGS, J, g.map);
}

template <int LT, int AR>
static void grp_sketch_t(bool omega, int fq, const GroupLaunch& g, hipStream_t st) {
  const JobSketch* J = static_cast<const JobSketch*>(g.jobs);
  if (omega) hipLaunchKernelGGL((k_sketch_grp<LT, AR, true, 1>), TTX_GRP_ARGS, J, g.map);
  else if (fq <= 1) hipLaunchKernelGGL((k_sketch_grp<LT, AR, false, 1>), TTX_GRP_ARGS, J, g.map);
  else if (fq == 2) hipLaunchKernelGGL((k_sketch_grp<LT, AR, false, 2>), TTX_GRP_ARGS, J, g.map);
  else hipLaunchKernelGGL((k_sketch_grp<LT, AR, false, 4>), TTX_GRP_ARGS, J, g.map);
}
template <int LT>
static void grp_sketch_ar(int ar, bool omega, int fq, const GroupLaunch& g, hipStream_t st) {
  switch (ar) {
    case AR_FP32: grp_sketch_t<LT, AR_FP32>(omega, fq, g, st); break;
    case AR_BF16: grp_sketch_t<LT, AR_BF16>(omega, fq, g, st); break;
    default: grp_sketch_t<LT, AR_BF16X3>(omega, fq, g, st); break;
  }
}
void grp_sketch(int ar, int lt, bool omega, int fq, const GroupLaunch& g, hipStream_t st) {
  if (lt == 1) grp_sketch_ar<1>(ar, omega, fq, g, st);
  else grp_sketch_ar<2>(ar, omega, fq, g, st);
}

template <int LT, bool GRAM>
static void grp_proj_t(int ar, const GroupLaunch& g, hipStream_t st) {
  const JobProj* J = static_cast<const JobProj*>(g.jobs);
  switch (ar) {
    case AR_FP32: hipLaunchKernelGGL((k_proj_grp<LT, AR_FP32, GRAM>), TTX_GRP_ARGS, J, g.map); break;
    case AR_BF16: hipLaunchKernelGGL((k_proj_grp<LT, AR_BF16, GRAM>), TTX_GRP_ARGS, J, g.map); break;
    case AR_BF16X4: hipLaunchKernelGGL((k_proj_grp<LT, AR_BF16X4, GRAM>), TTX_GRP_ARGS, J, g.map); break;
    default: hipLaunchKernelGGL((k_proj_grp<LT, AR_BF16X3, GRAM>), TTX_GRP_ARGS, J, g.map); break;
  }
}
void grp_proj_short(int ar, int lt, int nk, const GroupLaunch& g, hipStream_t st) {
  const JobProj* J = static_cast<const JobProj*>(g.jobs);
#define TTX_PS(LT, NK, AR) hipLaunchKernelGGL((k_proj_short_grp<LT, NK, AR>), TTX_GRP_ARGS, J, g.map)
#define TTX_PSN(LT, AR) { if (nk <= 1) TTX_PS(LT, 1, AR); else if (nk == 2) TTX_PS(LT, 2, AR); \
                          else if (nk == 3) TTX_PS(LT, 3, AR); else TTX_PS(LT, 4, AR); }
  if (ar == AR_FP32) { if (lt == 1) TTX_PSN(1, AR_FP32) else TTX_PSN(2, AR_FP32) }
  else { if (lt == 1) TTX_PSN(1, AR_BF16X4) else TTX_PSN(2, AR_BF16X4) }
#undef TTX_PSN
#undef TTX_PS
}

void grp_proj(int ar, int lt, bool gram, const GroupLaunch& g, hipStream_t st) {
  if (lt == 1) { if (gram) grp_proj_t<1, true>(ar, g, st); else grp_proj_t<1, false>(ar, g, st); }
  else { if (gram) grp_proj_t<2, true>(ar, g, st); else grp_proj_t<2, false>(ar, g, st); }
}
#undef TTX_GRP_ARGS

}  // namespace ttx
