// Streaming tall-skinny products over the snapshot increment dA = A_next - A_prev.
//
//   pass N:  P[i, c] = sum_j dA[i, j] * B[j, c]     (dA * V,  n x rb; reduction over the m columns)
//   pass T:  Q[j, c] = sum_i dA[i, j] * B[i, c]     (dA' * U, m x rb; reduction over the n rows)
//
// These are the n x m x r GEMMs of the reference's data sub-integrator
// `u .+= sign*left'*dy*right` (data_integrator.jl:14-16) as called by the K and
// L sub-steps (projector_splitting.jl:135,147,172,184; unconventional.jl:139,147;
// rank_adaptive_unconventional.jl:200,211) and the S sub-step
// (unconventional.jl:155; rank_adaptive_unconventional.jl:220).  dA is never
// materialised: both snapshots are streamed once per pass and subtracted in
// registers.  Snapshots are column-major (rows i contiguous).
//
// Every pass is split over its reduction range: split s writes the partial
// part[s][c][row] (FP64 for PASS_F64, FP32 otherwise); the FP64 row kernel sums
// the splits in a fixed order, so results are deterministic.
//
// Matrix-core instruction layouts used below (checked by lri_selftest_mfma):
//  f64 16x16x4:   A lane l -> A[l%16][l/16]; B lane l -> B[l/16][l%16];
//                 D lane l, reg v -> D[4*v+l/16][l%16]
//  bf16 32x32x16: A lane l, elem e -> A[l%32][8*(l/32)+e]; B -> B[8*(l/32)+e][l%32];
//                 D lane l, reg v -> D[8*(v/4)+4*(l/32)+(v%4)][l%32]
//  f32 32x32x2:   A lane l -> A[l%32][l/32]; B lane l -> B[l/32][l%32]; D as bf16.
// Only the pairing of A and B along k matters for a product, so the kernels
// assign k indices to (lane, element) freely as long as A and B agree.
#include "lri_common.h"
#include "lri_kernels.h"
#include "../../include/lri.h"

namespace lri {

constexpr int kPassThreads = 256;  // 4 waves

__device__ __forceinline__ f64x4_t mfma_f64(double a, double b, f64x4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16_t mfma_bf16(bf16x8_t a, bf16x8_t b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16_t mfma_f32(float a, float b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8_t as_bf16x8(u32x4_t v) { return __builtin_bit_cast(bf16x8_t, v); }

// ============================================================================
// FP64 passes (v_mfma_f64_16x16x4_f64); snapshots of any type decoded exactly
// ============================================================================

// pass N: wave tile = 16*TM rows x all rb columns, k-steps of 4 columns.
template <int TM, int CB, typename T>
__global__ __launch_bounds__(kPassThreads) void k_passN_f64(const T* __restrict__ an, const T* __restrict__ ap,
                                                            int64_t lda, int64_t n, int64_t m,
                                                            const double* __restrict__ B, int rb, int64_t jchunk,
                                                            double* __restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t i0 = ((int64_t)blockIdx.x * 4 + wave) * (16 * TM);
  if (i0 >= n) return;
  const int s = blockIdx.y;
  const int64_t jb = (int64_t)s * jchunk;
  const int64_t je = min(m, jb + jchunk);
  f64x4_t acc[TM][CB];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[t][c] = f64x4_t{0.0, 0.0, 0.0, 0.0};
  for (int64_t j = jb; j < je; j += 4) {
    const int64_t jj = j + lk;
    const bool jok = jj < je;
    double b[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) b[c] = jok ? B[jj * rb + c * 16 + lr] : 0.0;
#pragma unroll
    for (int t = 0; t < TM; ++t) {
      const int64_t ii = i0 + 16 * t + lr;
      double a = 0.0;
      if (jok && ii < n) {
        const int64_t off = jj * lda + ii;
        a = snap_f64(an, off) - snap_f64(ap, off);
      }
#pragma unroll
      for (int c = 0; c < CB; ++c) acc[t][c] = mfma_f64(a, b[c], acc[t][c]);
    }
  }
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t ii = i0 + 16 * t + 4 * v + lk;
        if (ii < n) part[((int64_t)s * rb + c * 16 + lr) * n + ii] = acc[t][c][v];
      }
}

// pass T: wave tile = 16*TJ columns of dA x all rb columns; each lane owns 4
// consecutive rows of its column per iteration (16 rows per iteration).
template <int TJ, int CB, typename T>
__global__ __launch_bounds__(kPassThreads) void k_passT_f64(const T* __restrict__ an, const T* __restrict__ ap,
                                                            int64_t lda, int64_t n, int64_t m,
                                                            const double* __restrict__ B, int rb, int64_t ichunk,
                                                            double* __restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t j0 = ((int64_t)blockIdx.x * 4 + wave) * (16 * TJ);
  if (j0 >= m) return;
  const int s = blockIdx.y;
  const int64_t ib = (int64_t)s * ichunk;
  const int64_t ie = min(n, ib + ichunk);
  f64x4_t acc[TJ][CB];
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[t][c] = f64x4_t{0.0, 0.0, 0.0, 0.0};
  for (int64_t i = ib; i < ie; i += 16) {
    double a[TJ][4];
#pragma unroll
    for (int t = 0; t < TJ; ++t) {
      const int64_t jj = j0 + 16 * t + lr;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t ii = i + 4 * lk + e;
        double v = 0.0;
        if (jj < m && ii < ie) {
          const int64_t off = jj * lda + ii;
          v = snap_f64(an, off) - snap_f64(ap, off);
        }
        a[t][e] = v;
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t ii = i + 4 * lk + e;
      double b[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) b[c] = ii < ie ? B[ii * rb + c * 16 + lr] : 0.0;
#pragma unroll
      for (int t = 0; t < TJ; ++t)
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[t][c] = mfma_f64(a[t][e], b[c], acc[t][c]);
    }
  }
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t jj = j0 + 16 * t + 4 * v + lk;
        if (jj < m) part[((int64_t)s * rb + c * 16 + lr) * m + jj] = acc[t][c][v];
      }
}

// ============================================================================
// FP32 passes (v_mfma_f32_32x32x2_f32) for F32 (or BF16) snapshots
// ============================================================================

template <typename T>
__device__ __forceinline__ float snap_f32(const T* p, int64_t k);
template <>
__device__ __forceinline__ float snap_f32<float>(const float* p, int64_t k) { return p[k]; }
template <>
__device__ __forceinline__ float snap_f32<uint16_t>(const uint16_t* p, int64_t k) { return bf16_to_f32(p[k]); }

// pass N: wave tile = 32*TM rows; k-steps of 2 columns.
template <int TM, int CB, typename T>
__global__ __launch_bounds__(kPassThreads) void k_passN_f32(const T* __restrict__ an, const T* __restrict__ ap,
                                                            int64_t lda, int64_t n, int64_t m,
                                                            const float* __restrict__ B, int rb, int64_t jchunk,
                                                            float* __restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 31, lk = lane >> 5;
  const int64_t i0 = ((int64_t)blockIdx.x * 4 + wave) * (32 * TM);
  if (i0 >= n) return;
  const int s = blockIdx.y;
  const int64_t jb = (int64_t)s * jchunk;
  const int64_t je = min(m, jb + jchunk);
  f32x16_t acc[TM][CB];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[t][c] = f32x16_t{};
  for (int64_t j = jb; j < je; j += 2) {
    const int64_t jj = j + lk;
    const bool jok = jj < je;
    float b[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) b[c] = jok ? B[jj * rb + c * 32 + lr] : 0.0f;
#pragma unroll
    for (int t = 0; t < TM; ++t) {
      const int64_t ii = i0 + 32 * t + lr;
      float a = 0.0f;
      if (jok && ii < n) {
        const int64_t off = jj * lda + ii;
        a = snap_f32(an, off) - snap_f32(ap, off);
      }
#pragma unroll
      for (int c = 0; c < CB; ++c) acc[t][c] = mfma_f32(a, b[c], acc[t][c]);
    }
  }
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t ii = i0 + 32 * t + 8 * (v / 4) + 4 * lk + (v % 4);
        if (ii < n) part[((int64_t)s * rb + c * 32 + lr) * n + ii] = acc[t][c][v];
      }
}

// pass T: wave tile = 32*TJ columns; each lane owns 4 consecutive rows per iteration (8 rows per iteration).
template <int TJ, int CB, typename T>
__global__ __launch_bounds__(kPassThreads) void k_passT_f32(const T* __restrict__ an, const T* __restrict__ ap,
                                                            int64_t lda, int64_t n, int64_t m,
                                                            const float* __restrict__ B, int rb, int64_t ichunk,
                                                            float* __restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 31, lk = lane >> 5;
  const int64_t j0 = ((int64_t)blockIdx.x * 4 + wave) * (32 * TJ);
  if (j0 >= m) return;
  const int s = blockIdx.y;
  const int64_t ib = (int64_t)s * ichunk;
  const int64_t ie = min(n, ib + ichunk);
  f32x16_t acc[TJ][CB];
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[t][c] = f32x16_t{};
  for (int64_t i = ib; i < ie; i += 8) {
    float a[TJ][4];
#pragma unroll
    for (int t = 0; t < TJ; ++t) {
      const int64_t jj = j0 + 32 * t + lr;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t ii = i + 4 * lk + e;
        float v = 0.0f;
        if (jj < m && ii < ie) {
          const int64_t off = jj * lda + ii;
          v = snap_f32(an, off) - snap_f32(ap, off);
        }
        a[t][e] = v;
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t ii = i + 4 * lk + e;
      float b[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) b[c] = ii < ie ? B[ii * rb + c * 32 + lr] : 0.0f;
#pragma unroll
      for (int t = 0; t < TJ; ++t)
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[t][c] = mfma_f32(a[t][e], b[c], acc[t][c]);
    }
  }
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t jj = j0 + 32 * t + 8 * (v / 4) + 4 * lk + (v % 4);
        if (jj < m) part[((int64_t)s * rb + c * 32 + lr) * m + jj] = acc[t][c][v];
      }
}

// ============================================================================
// BF16x3 passes (v_mfma_f32_32x32x16_bf16), BF16 snapshots.
// dA = A_next - A_prev is formed exactly in FP32, split dA = a_hi + a_lo into
// two BF16 values, the B panel is stored as b_hi + b_lo, and each product is
// a_hi*b_hi + a_hi*b_lo + a_lo*b_hi (the a_lo*b_lo term is below 2^-16
// relative), accumulated in FP32.
// ============================================================================

// Load 8 consecutive bf16 of a column (rows r0..r0+7, masked against rend)
// from both snapshots and return the FP32 differences.
__device__ __forceinline__ void load8_diff(const uint16_t* __restrict__ an, const uint16_t* __restrict__ ap,
                                           int64_t off, int64_t r0, int64_t rend, bool colok, bool vec,
                                           float (&d)[8]) {
  if (colok && vec && r0 + 8 <= rend) {
    const u32x4_t x = *reinterpret_cast<const u32x4_t*>(an + off);
    const u32x4_t y = *reinterpret_cast<const u32x4_t*>(ap + off);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      d[2 * q] = __uint_as_float(x[q] << 16) - __uint_as_float(y[q] << 16);
      d[2 * q + 1] = __uint_as_float(x[q] & 0xFFFF0000u) - __uint_as_float(y[q] & 0xFFFF0000u);
    }
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e)
      d[e] = (colok && r0 + e < rend) ? bf16_to_f32(an[off + e]) - bf16_to_f32(ap[off + e]) : 0.0f;
  }
}

__device__ __forceinline__ void split8(const float (&d)[8], u32x4_t& hi, u32x4_t& lo) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t h0, l0, h1, l1;
    split_bf16(d[2 * q], h0, l0);
    split_bf16(d[2 * q + 1], h1, l1);
    hi[q] = h0 | (h1 << 16);
    lo[q] = l0 | (l1 << 16);
  }
}

// pass T: wave tile = 32*TJ columns of dA; each lane owns 32 consecutive rows
// of its column per iteration (64 rows per iteration, 4 k-steps of 16).
template <int TJ, int CB>
__global__ __launch_bounds__(kPassThreads) void k_passT_bf16x3(const uint16_t* __restrict__ an,
                                                               const uint16_t* __restrict__ ap, int64_t lda,
                                                               int64_t n, int64_t m, const uint16_t* __restrict__ Bhi,
                                                               const uint16_t* __restrict__ Blo, int64_t ldb,
                                                               int rb, int64_t ichunk, float* __restrict__ part,
                                                               int vec) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 31, h = lane >> 5;
  const int64_t j0 = ((int64_t)blockIdx.x * 4 + wave) * (32 * TJ);
  if (j0 >= m) return;
  const int s = blockIdx.y;
  const int64_t ib = (int64_t)s * ichunk;
  const int64_t ie = min(n, ib + ichunk);
  f32x16_t acc[TJ][CB];
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[t][c] = f32x16_t{};
  for (int64_t i = ib; i < ie; i += 64) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t r0 = i + 32 * h + 8 * q;
      u32x4_t bh[CB], bl[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        const int64_t boff = (int64_t)(c * 32 + lc) * ldb + r0;
        bh[c] = *reinterpret_cast<const u32x4_t*>(Bhi + boff);
        bl[c] = *reinterpret_cast<const u32x4_t*>(Blo + boff);
      }
#pragma unroll
      for (int t = 0; t < TJ; ++t) {
        const int64_t jj = j0 + 32 * t + lc;
        float d[8];
        load8_diff(an, ap, jj * lda + r0, r0, ie, jj < m, vec, d);
        u32x4_t ah, al;
        split8(d, ah, al);
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          acc[t][c] = mfma_bf16(as_bf16x8(ah), as_bf16x8(bh[c]), acc[t][c]);
          acc[t][c] = mfma_bf16(as_bf16x8(ah), as_bf16x8(bl[c]), acc[t][c]);
          acc[t][c] = mfma_bf16(as_bf16x8(al), as_bf16x8(bh[c]), acc[t][c]);
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t jj = j0 + 32 * t + 8 * (v / 4) + 4 * h + (v % 4);
        if (jj < m) part[((int64_t)s * rb + c * 32 + lc) * m + jj] = acc[t][c][v];
      }
}

// pass N: the operand needs 8 consecutive columns per lane, but memory runs
// along rows, so each wave stages a (32*TM rows x 16 columns) hi/lo tile
// through its own LDS region: coalesced 16-byte loads along the columns,
// stored column-major, then read back as 8 column-consecutive bf16 per lane.
constexpr int kNPad = 4;  // LDS column stride = rows + 4 (bf16) keeps the two lane halves on different banks

template <int TM, int CB>
__global__ __launch_bounds__(kPassThreads) void k_passN_bf16x3(const uint16_t* __restrict__ an,
                                                               const uint16_t* __restrict__ ap, int64_t lda,
                                                               int64_t n, int64_t m, const uint16_t* __restrict__ Bhi,
                                                               const uint16_t* __restrict__ Blo, int64_t ldb,
                                                               int rb, int64_t jchunk, float* __restrict__ part,
                                                               int vec) {
  constexpr int ROWS = 32 * TM;
  constexpr int LDC = ROWS + kNPad;  // column stride in the LDS tile (bf16 elements)
  __shared__ uint16_t tile[4][2][16 * LDC];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 31, h = lane >> 5;
  const int64_t i0 = ((int64_t)blockIdx.x * 4 + wave) * ROWS;
  const bool active = i0 < n;
  const int s = blockIdx.y;
  const int64_t jb = (int64_t)s * jchunk;
  const int64_t je = min(m, jb + jchunk);
  uint16_t* th = tile[wave][0];
  uint16_t* tl = tile[wave][1];
  f32x16_t acc[TM][CB];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[t][c] = f32x16_t{};
  for (int64_t j = jb; j < je; j += 16) {
    // stage: TM 8-row chunks per lane; chunk b -> column b / (4*TM), row octet b % (4*TM)
    if (active) {
#pragma unroll
      for (int q = 0; q < TM; ++q) {
        const int b = q * 64 + lane;
        const int col = b / (4 * TM), oct = b % (4 * TM);
        const int64_t jj = j + col;
        const int64_t r0 = i0 + 8 * oct;
        float d[8];
        load8_diff(an, ap, jj * lda + r0, r0, n, jj < je, vec, d);
        u32x4_t hv, lv;
        split8(d, hv, lv);
        uint32_t* ph = reinterpret_cast<uint32_t*>(th + col * LDC + 8 * oct);
        uint32_t* pl = reinterpret_cast<uint32_t*>(tl + col * LDC + 8 * oct);
        // LDC*2 bytes is 8-byte aligned: two 8-byte stores per 16-byte chunk
        reinterpret_cast<uint2*>(ph)[0] = uint2{hv[0], hv[1]};
        reinterpret_cast<uint2*>(ph)[1] = uint2{hv[2], hv[3]};
        reinterpret_cast<uint2*>(pl)[0] = uint2{lv[0], lv[1]};
        reinterpret_cast<uint2*>(pl)[1] = uint2{lv[2], lv[3]};
      }
    }
    __syncthreads();
    if (active) {
      u32x4_t bh[CB], bl[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        const int64_t boff = (int64_t)(c * 32 + lc) * ldb + j + 8 * h;
        bh[c] = *reinterpret_cast<const u32x4_t*>(Bhi + boff);
        bl[c] = *reinterpret_cast<const u32x4_t*>(Blo + boff);
      }
#pragma unroll
      for (int t = 0; t < TM; ++t) {
        const int row = 32 * t + lc;
        u32x4_t ah, al;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int c0 = 8 * h + 2 * q;
          ah[q] = uint32_t(th[c0 * LDC + row]) | (uint32_t(th[(c0 + 1) * LDC + row]) << 16);
          al[q] = uint32_t(tl[c0 * LDC + row]) | (uint32_t(tl[(c0 + 1) * LDC + row]) << 16);
        }
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          acc[t][c] = mfma_bf16(as_bf16x8(ah), as_bf16x8(bh[c]), acc[t][c]);
          acc[t][c] = mfma_bf16(as_bf16x8(ah), as_bf16x8(bl[c]), acc[t][c]);
          acc[t][c] = mfma_bf16(as_bf16x8(al), as_bf16x8(bh[c]), acc[t][c]);
        }
      }
    }
    __syncthreads();
  }
  if (!active) return;
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t ii = i0 + 32 * t + 8 * (v / 4) + 4 * h + (v % 4);
        if (ii < n) part[((int64_t)s * rb + c * 32 + lc) * n + ii] = acc[t][c][v];
      }
}

// ============================================================================
// planning and launch
// ============================================================================

// Tile height per wave (rows of dA for pass N, columns of dA for pass T), in
// units of the instruction's M (16 for FP64, 32 otherwise).  Accumulators are
// tile x (rb / N) blocks, kept at <= 8 (FP64) or 4 (FP32) blocks per wave.
static int tile_units(int mode, bool transposed, int rb) {
  if (mode == PASS_F64) {
    int cb = rb / 16;
    return cb <= 2 ? 4 : (cb <= 4 ? 2 : 1);
  }
  int cb = rb / 32;
  if (mode == PASS_F32) return cb <= 1 ? 2 : 1;
  if (transposed) return cb <= 1 ? 2 : 1;
  return cb <= 1 ? 4 : (cb == 2 ? 2 : 1);
}

static int wg_span(int mode, bool transposed, int rb) {
  return 4 * (mode == PASS_F64 ? 16 : 32) * tile_units(mode, transposed, rb);
}

int64_t plan_pass(int mode, bool transposed, int64_t out_rows, int64_t red_len, int rb, int* splits,
                  int64_t* chunk) {
  const int span = wg_span(mode, transposed, rb);
  const int64_t wgs = (out_rows + span - 1) / span;
  const int64_t quantum = 64;  // reduction chunks are multiples of 64 (alignment of the bf16 tiles)
  // aim for ~2048 workgroups, chunks of at least 256 along the reduction
  int64_t want = (2048 + wgs - 1) / wgs;
  int64_t maxs = (red_len + 255) / 256;
  if (want > maxs) want = maxs;
  if (want < 1) want = 1;
  int64_t ch = (red_len + want - 1) / want;
  ch = (ch + quantum - 1) / quantum * quantum;
  int64_t s = (red_len + ch - 1) / ch;
  *splits = (int)s;
  *chunk = ch;
  return s * rb * out_rows;
}

static bool vec_ok(const PassArgs& a) {
  return (a.lda % 8 == 0) && ((uintptr_t)a.anext % 16 == 0) && ((uintptr_t)a.aprev % 16 == 0);
}

template <int TM, int CB>
static hipError_t run_pass(int mode, bool transposed, const PassArgs& a, hipStream_t st) {
  const int64_t out_rows = transposed ? a.m : a.n;
  const int span = 4 * (mode == PASS_F64 ? 16 : 32) * TM;
  dim3 grid((unsigned)((out_rows + span - 1) / span), (unsigned)a.splits);
  dim3 blk(kPassThreads);
  if (mode == PASS_F64) {
    const double* B = (const double*)a.B;
    double* P = (double*)a.part;
#define LRI_F64_LAUNCH(T)                                                                                    \
  if (transposed)                                                                                            \
    hipLaunchKernelGGL((k_passT_f64<TM, CB, T>), grid, blk, 0, st, (const T*)a.anext, (const T*)a.aprev,      \
                       a.lda, a.n, a.m, B, a.rb, a.chunk, P);                                                \
  else                                                                                                       \
    hipLaunchKernelGGL((k_passN_f64<TM, CB, T>), grid, blk, 0, st, (const T*)a.anext, (const T*)a.aprev,      \
                       a.lda, a.n, a.m, B, a.rb, a.chunk, P);
    if (a.dtype == LRI_F64) {
      LRI_F64_LAUNCH(double)
    } else if (a.dtype == LRI_F32) {
      LRI_F64_LAUNCH(float)
    } else {
      LRI_F64_LAUNCH(uint16_t)
    }
#undef LRI_F64_LAUNCH
  } else if (mode == PASS_F32) {
    const float* B = (const float*)a.B;
    float* P = (float*)a.part;
#define LRI_F32_LAUNCH(T)                                                                                    \
  if (transposed)                                                                                            \
    hipLaunchKernelGGL((k_passT_f32<TM, CB, T>), grid, blk, 0, st, (const T*)a.anext, (const T*)a.aprev,      \
                       a.lda, a.n, a.m, B, a.rb, a.chunk, P);                                                \
  else                                                                                                       \
    hipLaunchKernelGGL((k_passN_f32<TM, CB, T>), grid, blk, 0, st, (const T*)a.anext, (const T*)a.aprev,      \
                       a.lda, a.n, a.m, B, a.rb, a.chunk, P);
    if (a.dtype == LRI_F32) {
      LRI_F32_LAUNCH(float)
    } else {
      LRI_F32_LAUNCH(uint16_t)
    }
#undef LRI_F32_LAUNCH
  } else {
    const uint16_t* an = (const uint16_t*)a.anext;
    const uint16_t* ap = (const uint16_t*)a.aprev;
    const int v = vec_ok(a) ? 1 : 0;
    if (transposed)
      hipLaunchKernelGGL((k_passT_bf16x3<TM, CB>), grid, blk, 0, st, an, ap, a.lda, a.n, a.m,
                         (const uint16_t*)a.B, (const uint16_t*)a.B2, a.ldb, a.rb, a.chunk, (float*)a.part, v);
    else
      hipLaunchKernelGGL((k_passN_bf16x3<TM, CB>), grid, blk, 0, st, an, ap, a.lda, a.n, a.m,
                         (const uint16_t*)a.B, (const uint16_t*)a.B2, a.ldb, a.rb, a.chunk, (float*)a.part, v);
  }
  return hipGetLastError();
}

static hipError_t launch_pass(int mode, bool transposed, const PassArgs& a, hipStream_t st) {
  if (mode == PASS_BF16X3 && a.dtype != LRI_BF16) return hipErrorInvalidValue;
  if (mode == PASS_F32 && a.dtype == LRI_F64) return hipErrorInvalidValue;
  const int tu = tile_units(mode, transposed, a.rb);
  if (mode == PASS_F64) {
    switch (a.rb) {
      case 16: return run_pass<4, 1>(mode, transposed, a, st);
      case 32: return run_pass<4, 2>(mode, transposed, a, st);
      case 48: return run_pass<2, 3>(mode, transposed, a, st);
      case 64: return run_pass<2, 4>(mode, transposed, a, st);
      case 128: return run_pass<1, 8>(mode, transposed, a, st);
      default: return hipErrorInvalidValue;
    }
  }
  switch (a.rb) {
    case 32: return tu == 4 ? run_pass<4, 1>(mode, transposed, a, st) : run_pass<2, 1>(mode, transposed, a, st);
    case 64: return tu == 2 ? run_pass<2, 2>(mode, transposed, a, st) : run_pass<1, 2>(mode, transposed, a, st);
    case 128: return run_pass<1, 4>(mode, transposed, a, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_passN(int mode, const PassArgs& a, hipStream_t st) { return launch_pass(mode, false, a, st); }
hipError_t launch_passT(int mode, const PassArgs& a, hipStream_t st) { return launch_pass(mode, true, a, st); }

// ============================================================================
// lane-layout self test: one wave, one instruction, raw results per lane
// ============================================================================

__global__ void k_selftest(int which, const void* a, const void* b, void* out) {
  const int l = threadIdx.x;
  if (which == 0) {
    f64x4_t d = mfma_f64(((const double*)a)[l], ((const double*)b)[l], f64x4_t{0.0, 0.0, 0.0, 0.0});
    for (int v = 0; v < 4; ++v) ((double*)out)[l * 4 + v] = d[v];
  } else if (which == 1) {
    const u32x4_t av = reinterpret_cast<const u32x4_t*>(a)[l];
    const u32x4_t bv = reinterpret_cast<const u32x4_t*>(b)[l];
    f32x16_t d = mfma_bf16(as_bf16x8(av), as_bf16x8(bv), f32x16_t{});
    for (int v = 0; v < 16; ++v) ((float*)out)[l * 16 + v] = d[v];
  } else {
    f32x16_t d = mfma_f32(((const float*)a)[l], ((const float*)b)[l], f32x16_t{});
    for (int v = 0; v < 16; ++v) ((float*)out)[l * 16 + v] = d[v];
  }
}

hipError_t selftest_mfma(int which, const void* a, const void* b, void* out) {
  size_t in_bytes = which == 0 ? 64 * 8 : (which == 1 ? 64 * 16 : 64 * 4);
  size_t out_bytes = which == 0 ? 256 * 8 : 1024 * 4;
  void *da = nullptr, *db = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&da, in_bytes);
  if (e == hipSuccess) e = hipMalloc(&db, in_bytes);
  if (e == hipSuccess) e = hipMalloc(&dout, out_bytes);
  if (e == hipSuccess) e = hipMemcpy(da, a, in_bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, b, in_bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_selftest, dim3(1), dim3(64), 0, 0, which, da, db, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, dout, out_bytes, hipMemcpyDeviceToHost);
  if (da) (void)hipFree(da);
  if (db) (void)hipFree(db);
  if (dout) (void)hipFree(dout);
  return e;
}

}  // namespace lri
