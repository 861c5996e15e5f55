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
// Work decomposition shared by all passes
//
// grid.x = output tiles (rows of dA for pass N, columns for pass T),
// grid.y = splits of the reduction range.  The 4 waves of a workgroup share
// the output tile; wave w streams the w-th contiguous quarter of the split's
// range, prefetching the next iteration's operands while the current one is
// multiplied.  The four accumulators are summed through LDS in a fixed order
// (deterministic) and wave 0 writes the split's partial.
// ============================================================================

struct WaveRange {
  int64_t b, e;
};

// [lo, hi) cut into 4 contiguous pieces whose lengths are multiples of q (the last one is shorter).
__device__ __forceinline__ WaveRange wave_range(int64_t lo, int64_t hi, int wave, int64_t q) {
  const int64_t per = ((hi - lo + 3) / 4 + q - 1) / q * q;
  const int64_t b = min(hi, lo + wave * per);
  return {b, min(hi, b + per)};
}

// Sum the accumulators of waves 1..3 into wave 0 (LDS red: 3 * NV * 64 values).
template <typename T, int NV>
__device__ __forceinline__ void wg_sum(T (&acc)[NV], T* red, int wave, int lane) {
  __syncthreads();
  if (wave > 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) red[((wave - 1) * NV + v) * 64 + lane] = acc[v];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v)
      acc[v] = ((acc[v] + red[v * 64 + lane]) + red[(NV + v) * 64 + lane]) + red[(2 * NV + v) * 64 + lane];
  }
}

// Order LDS stores of one wave before its later loads (and vice versa).
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ============================================================================
// FP64 passes (v_mfma_f64_16x16x4_f64); snapshots of any type decoded exactly
// ============================================================================

// N consecutive elements (N = 1, 2 or 4) of a column (rows r0 .. r0+N-1,
// masked against rend), decoded exactly to FP64; vector loads when aligned and
// in range.
template <int N, typename T>
__device__ __forceinline__ void load_rows_f64(const T* __restrict__ p, int64_t off, int64_t r0, int64_t rend,
                                              bool ok, bool vec, double (&v)[N]) {
  if (ok && vec && r0 + N <= rend) {
    if constexpr (sizeof(T) == 8 && N == 4) {
      const double2 x = *reinterpret_cast<const double2*>(p + off);
      const double2 y = *reinterpret_cast<const double2*>(p + off + 2);
      v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
      return;
    } else if constexpr (sizeof(T) == 8 && N == 2) {
      const double2 x = *reinterpret_cast<const double2*>(p + off);
      v[0] = x.x; v[1] = x.y;
      return;
    } else if constexpr (sizeof(T) == 4 && N == 4) {
      const float4 x = *reinterpret_cast<const float4*>(p + off);
      v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
      return;
    } else if constexpr (sizeof(T) == 4 && N == 2) {
      const float2 x = *reinterpret_cast<const float2*>(p + off);
      v[0] = x.x; v[1] = x.y;
      return;
    } else if constexpr (sizeof(T) == 2 && N == 4) {
      const uint2 x = *reinterpret_cast<const uint2*>(p + off);
      v[0] = __uint_as_float(x.x << 16); v[1] = __uint_as_float(x.x & 0xFFFF0000u);
      v[2] = __uint_as_float(x.y << 16); v[3] = __uint_as_float(x.y & 0xFFFF0000u);
      return;
    } else if constexpr (sizeof(T) == 2 && N == 2) {
      const uint32_t x = *reinterpret_cast<const uint32_t*>(p + off);
      v[0] = __uint_as_float(x << 16); v[1] = __uint_as_float(x & 0xFFFF0000u);
      return;
    }
  }
#pragma unroll
  for (int e = 0; e < N; ++e) v[e] = (ok && r0 + e < rend) ? snap_f64(p, off + e) : 0.0;
}

// pass N: tile = 16*TM rows.  Lane (lr = l%16, lk = l/16) owns rows
// i0 + TM*lr + e (e < TM) and feeds them to TM instructions (row block e,
// instruction row lr), so each lane streams TM consecutive rows of a column per
// load; an iteration covers 8 columns (2 k-steps of 4).
template <int TM, int CB, typename T>
__global__ __launch_bounds__(kPassThreads) void k_passN_f64(const T* __restrict__ an, const T* __restrict__ ap,
                                                            int64_t lda, int64_t n, int64_t m,
                                                            const double* __restrict__ B, int rb, int64_t jchunk,
                                                            double* __restrict__ part, int vec) {
  constexpr int NV = TM * CB * 4;
  __shared__ double red[3 * NV * 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t i0 = (int64_t)blockIdx.x * (16 * TM);
  const int64_t r0 = i0 + TM * lr;
  const int s = blockIdx.y;
  const int64_t jb = (int64_t)s * jchunk;
  const int64_t je = min(m, jb + jchunk);
  const WaveRange R = wave_range(jb, je, wave, 8);
  const bool v = vec != 0;
  f64x4_t acc[TM][CB];
#pragma unroll
  for (int e = 0; e < TM; ++e)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[e][c] = f64x4_t{0.0, 0.0, 0.0, 0.0};
  double xa[2][TM], xp[2][TM], b[2][CB], nxa[2][TM], nxp[2][TM], nb[2][CB];
  auto load = [&](int64_t j, double (&A1)[2][TM], double (&A0)[2][TM], double (&Bv)[2][CB]) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int64_t jj = j + 4 * u + lk;
      const bool jok = jj < je;
      load_rows_f64<TM>(an, jj * lda + r0, r0, n, jok, v, A1[u]);
      load_rows_f64<TM>(ap, jj * lda + r0, r0, n, jok && ap != nullptr, v, A0[u]);  // no ap: A itself
#pragma unroll
      for (int c = 0; c < CB; ++c) Bv[u][c] = jok ? B[jj * rb + c * 16 + lr] : 0.0;
    }
  };
  if (R.b < R.e) load(R.b, xa, xp, b);
  for (int64_t j = R.b; j < R.e; j += 8) {
    const bool more = j + 8 < R.e;
    if (more) load(j + 8, nxa, nxp, nb);
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int e = 0; e < TM; ++e) {
        const double d = xa[u][e] - xp[u][e];
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[e][c] = mfma_f64(d, b[u][c], acc[e][c]);
      }
    if (more) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
#pragma unroll
        for (int e = 0; e < TM; ++e) {
          xa[u][e] = nxa[u][e];
          xp[u][e] = nxp[u][e];
        }
#pragma unroll
        for (int c = 0; c < CB; ++c) b[u][c] = nb[u][c];
      }
    }
  }
  double flat[NV];
#pragma unroll
  for (int e = 0; e < TM; ++e)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int w = 0; w < 4; ++w) flat[(e * CB + c) * 4 + w] = acc[e][c][w];
  wg_sum<double, NV>(flat, red, wave, lane);
  if (wave != 0) return;
  // instruction row 4w + lk of row block e is dA row i0 + TM*(4w + lk) + e
#pragma unroll
  for (int e = 0; e < TM; ++e)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const int64_t ii = i0 + TM * (4 * w + lk) + e;
        if (ii < n) part[((int64_t)s * rb + c * 16 + lr) * n + ii] = flat[(e * CB + c) * 4 + w];
      }
}

// pass T: tile = 16*TJ columns of dA; an iteration covers 16 rows, each lane
// owning 4 consecutive rows of its column (one vector load per snapshot).
template <int TJ, int CB, typename T>
__global__ __launch_bounds__(kPassThreads) void k_passT_f64(const T* __restrict__ an, const T* __restrict__ ap,
                                                            int64_t lda, int64_t n, int64_t m,
                                                            const double* __restrict__ B, int rb, int64_t ichunk,
                                                            double* __restrict__ part, int vec) {
  constexpr int NV = TJ * CB * 4;
  __shared__ double red[3 * NV * 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t j0 = (int64_t)blockIdx.x * (16 * TJ);
  const int s = blockIdx.y;
  const int64_t ib = (int64_t)s * ichunk;
  const int64_t ie = min(n, ib + ichunk);
  const WaveRange R = wave_range(ib, ie, wave, 16);
  const bool v = vec != 0;
  f64x4_t acc[TJ][CB];
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[t][c] = f64x4_t{0.0, 0.0, 0.0, 0.0};
  double xa[TJ][4], xp[TJ][4], b[4][CB], nxa[TJ][4], nxp[TJ][4], nb[4][CB];
  auto load = [&](int64_t i, double (&A1)[TJ][4], double (&A0)[TJ][4], double (&Bv)[4][CB]) {
    const int64_t r0 = i + 4 * lk;
#pragma unroll
    for (int t = 0; t < TJ; ++t) {
      const int64_t jj = j0 + 16 * t + lr;
      load_rows_f64<4>(an, jj * lda + r0, r0, ie, jj < m, v, A1[t]);
      load_rows_f64<4>(ap, jj * lda + r0, r0, ie, jj < m && ap != nullptr, v, A0[t]);  // no ap: A itself
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t ii = r0 + e;
#pragma unroll
      for (int c = 0; c < CB; ++c) Bv[e][c] = ii < ie ? B[ii * rb + c * 16 + lr] : 0.0;
    }
  };
  if (R.b < R.e) load(R.b, xa, xp, b);
  for (int64_t i = R.b; i < R.e; i += 16) {
    const bool more = i + 16 < R.e;
    if (more) load(i + 16, nxa, nxp, nb);
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int t = 0; t < TJ; ++t) {
        const double d = xa[t][e] - xp[t][e];
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[t][c] = mfma_f64(d, b[e][c], acc[t][c]);
      }
    if (more) {
#pragma unroll
      for (int t = 0; t < TJ; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          xa[t][e] = nxa[t][e];
          xp[t][e] = nxp[t][e];
        }
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int c = 0; c < CB; ++c) b[e][c] = nb[e][c];
    }
  }
  double flat[NV];
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int w = 0; w < 4; ++w) flat[(t * CB + c) * 4 + w] = acc[t][c][w];
  wg_sum<double, NV>(flat, red, wave, lane);
  if (wave != 0) return;
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const int64_t jj = j0 + 16 * t + 4 * w + lk;
        if (jj < m) part[((int64_t)s * rb + c * 16 + lr) * m + jj] = flat[(t * CB + c) * 4 + w];
      }
}

// ============================================================================
// FP32 passes (v_mfma_f32_32x32x2_f32) for F32 snapshots
// ============================================================================

typedef __attribute__((ext_vector_type(2))) float f32x2v_t;

// TM consecutive floats of a column (rows r0 .. r0+TM-1 masked against rend) as
// one vector load when aligned and in range.
template <int TM>
__device__ __forceinline__ void load_rows_f32(const float* __restrict__ p, int64_t off, int64_t r0, int64_t rend,
                                              bool ok, bool vec, float (&v)[TM]) {
  if (ok && vec && r0 + TM <= rend) {
    if constexpr (TM == 4) {
      const f32x4_t x = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p + off));
      v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; v[3] = x[3];
      return;
    } else if constexpr (TM == 2) {
      const f32x2v_t x = __builtin_nontemporal_load(reinterpret_cast<const f32x2v_t*>(p + off));
      v[0] = x[0]; v[1] = x[1];
      return;
    }
  }
#pragma unroll
  for (int e = 0; e < TM; ++e) v[e] = (ok && r0 + e < rend) ? p[off + e] : 0.0f;
}

// pass N: tile = 32*TM rows.  Lane (lr = l%32, h = l/32) owns the TM consecutive
// rows i0 + TM*lr .. i0 + TM*lr + TM-1 of a column (one 16-byte load for TM = 4)
// and feeds row i0 + TM*lr + t to instruction t, so accumulator t holds the
// output rows i0 + TM*d + t (d = 0..31).  A wave load instruction covers two
// columns of 128*TM contiguous bytes.  An iteration covers 8 columns (4 k-steps
// of 2); B (V) is row-major [j][rb], L2-resident.
template <int TM, int CB>
__global__ __launch_bounds__(kPassThreads) void k_passN_f32(const float* __restrict__ an,
                                                            const float* __restrict__ ap, int64_t lda, int64_t n,
                                                            int64_t m, const float* __restrict__ B, int rb,
                                                            int64_t jchunk, float* __restrict__ part, int vec) {
  constexpr int NV = TM * CB * 16;
  __shared__ float red[3 * NV * 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 31, h = lane >> 5;
  const int64_t i0 = (int64_t)blockIdx.x * (32 * TM);
  const int64_t r0 = i0 + TM * lr;
  const int s = blockIdx.y;
  const int64_t jb = (int64_t)s * jchunk;
  const int64_t je = min(m, jb + jchunk);
  const WaveRange R = wave_range(jb, je, wave, 8);
  const bool v = vec != 0;
  f32x16_t acc[TM][CB];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[t][c] = f32x16_t{};
  float xa[4][TM], xp[4][TM], b[4][CB], nxa[4][TM], nxp[4][TM], nb[4][CB];
  auto load = [&](int64_t j, float (&A1)[4][TM], float (&A0)[4][TM], float (&Bv)[4][CB]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t jj = j + 2 * u + h;
      const bool jok = jj < je;
      load_rows_f32<TM>(an, jj * lda + r0, r0, n, jok, v, A1[u]);
      load_rows_f32<TM>(ap, jj * lda + r0, r0, n, jok && ap != nullptr, v, A0[u]);  // no ap: A itself
#pragma unroll
      for (int c = 0; c < CB; ++c) Bv[u][c] = jok ? B[jj * rb + c * 32 + lr] : 0.0f;
    }
  };
  if (R.b < R.e) load(R.b, xa, xp, b);
  for (int64_t j = R.b; j < R.e; j += 8) {
    const bool more = j + 8 < R.e;
    if (more) load(j + 8, nxa, nxp, nb);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int t = 0; t < TM; ++t) {
        const float d = xa[u][t] - xp[u][t];
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[t][c] = mfma_f32(d, b[u][c], acc[t][c]);
      }
    if (more) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int t = 0; t < TM; ++t) {
          xa[u][t] = nxa[u][t];
          xp[u][t] = nxp[u][t];
        }
#pragma unroll
        for (int c = 0; c < CB; ++c) b[u][c] = nb[u][c];
      }
    }
  }
  float flat[NV];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int w = 0; w < 16; ++w) flat[(t * CB + c) * 16 + w] = acc[t][c][w];
  wg_sum<float, NV>(flat, red, wave, lane);
  if (wave != 0) return;
  // instruction row d = 8 (w / 4) + 4 h + w % 4 of accumulator t is output row i0 + TM d + t
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int w = 0; w < 16; ++w) {
        const int64_t ii = i0 + TM * (8 * (w / 4) + 4 * h + (w % 4)) + t;
        if (ii < n) part[((int64_t)s * rb + c * 32 + lr) * n + ii] = flat[(t * CB + c) * 16 + w];
      }
}

// pass T, wide form: a workgroup covers 128 columns of dA (wave w the columns
// 32w .. 32w+31), all four waves over the same rows, 32 rows per iteration.
// Each wave reads its 32 x 32 tile of each snapshot with 8 lanes per column (128
// contiguous bytes, 8 columns per wave instruction), stages the FP32 differences
// column-major in its LDS region (pitch 36 floats) and reads back, as lane
// (column lc, half h), the rows 16h .. 16h+15 of its column for k-steps
// e = 0..15.  The B block of the 32 rows (pass-T order, see pk_f32_offset: the
// 16 rows of (half h, column c) contiguous) is loaded once per workgroup into
// double-buffered LDS, prefetched into registers one block ahead.  Every wave
// owns its 32 output rows: no cross-wave reduction.
constexpr int kF32Pitch = 36;

template <int CB>
__global__ __launch_bounds__(kPassThreads) void k_passT_f32_wide(const float* __restrict__ an,
                                                                 const float* __restrict__ ap, int64_t lda,
                                                                 int64_t n, int64_t m, const float* __restrict__ B,
                                                                 int rb, int64_t ichunk, float* __restrict__ part,
                                                                 int vec) {
  constexpr int BQ = 32 * 32 * CB / 4 / kPassThreads;  // 16-byte pieces of a B block per thread
  __shared__ __attribute__((aligned(16))) float sA[4][32 * kF32Pitch];
  __shared__ f32x4_t sB[2][BQ * kPassThreads];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int lc = lane & 31, h = lane >> 5;
  const int64_t j0 = (int64_t)blockIdx.x * 128 + 32 * wave;
  const int s = blockIdx.y;
  const int64_t ib = (int64_t)s * ichunk;
  const int64_t ie = min(n, ib + ichunk);
  const bool v = vec != 0;
  float* tA = sA[wave];
  f32x16_t acc[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) acc[c] = f32x16_t{};
  // load lane mapping: instruction p covers columns 8p .. 8p+7; lane -> column 8p + lane/8, rows 4*(lane%8)
  const int lcol = lane >> 3, loct = lane & 7;
  float xn[4][4], xp[4][4], nxn[4][4], nxp[4][4];
  auto load = [&](int64_t i, float (&A1)[4][4], float (&A0)[4][4]) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int64_t jj = j0 + 8 * p + lcol;
      const int64_t r0 = i + 4 * loct;
      load_rows_f32<4>(an, jj * lda + r0, r0, ie, jj < m, v, A1[p]);
      load_rows_f32<4>(ap, jj * lda + r0, r0, ie, jj < m && ap != nullptr, v, A0[p]);  // no ap: A itself
    }
  };
  auto loadB = [&](int64_t i, f32x4_t (&g)[BQ]) {
    const f32x4_t* src = reinterpret_cast<const f32x4_t*>(B + (i >> 5) * 32 * rb);
#pragma unroll
    for (int q = 0; q < BQ; ++q) g[q] = src[q * kPassThreads + t];
  };
  f32x4_t g[BQ];
  int buf = 0;
  if (ib < ie) {
    load(ib, xn, xp);
    loadB(ib, g);
#pragma unroll
    for (int q = 0; q < BQ; ++q) sB[0][q * kPassThreads + t] = g[q];
  }
  __syncthreads();
  for (int64_t i = ib; i < ie; i += 32) {
    const bool more = i + 32 < ie;
    if (more) {
      load(i + 32, nxn, nxp);
      loadB(i + 32, g);
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int col = 8 * p + lcol;
      f32x4_t d;
#pragma unroll
      for (int e = 0; e < 4; ++e) d[e] = xn[p][e] - xp[p][e];
      *reinterpret_cast<f32x4_t*>(tA + col * kF32Pitch + 4 * loct) = d;
    }
    wave_lds_fence();
    f32x4_t a4[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) a4[q] = *reinterpret_cast<const f32x4_t*>(tA + lc * kF32Pitch + 16 * h + 4 * q);
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const f32x4_t* bp = reinterpret_cast<const f32x4_t*>(&sB[buf][0]) + (h * rb + c * 32 + lc) * 4;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const f32x4_t b4 = bp[q];
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[c] = mfma_f32(a4[q][e], b4[e], acc[c]);
      }
    }
    if (more) {
#pragma unroll
      for (int q = 0; q < BQ; ++q) sB[buf ^ 1][q * kPassThreads + t] = g[q];
    }
    __syncthreads();  // sB[buf] read by all waves, sB[buf ^ 1] written; the A staging is free again
    if (more) {
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          xn[p][e] = nxn[p][e];
          xp[p][e] = nxp[p][e];
        }
    }
    buf ^= 1;
  }
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      const int64_t jj = j0 + 8 * (w / 4) + 4 * h + (w % 4);
      if (jj < m) part[((int64_t)s * rb + c * 32 + lc) * m + jj] = acc[c][w];
    }
}

// ============================================================================
// BF16x3 passes (v_mfma_f32_32x32x16_bf16), BF16 snapshots.
// dA = A_next - A_prev is formed exactly in FP32, split dA = a_hi + a_lo into
// two BF16 values (v_cvt_pk_bf16_f32, round to nearest even); the B panel is
// stored as b_hi + b_lo, and each product is a_hi*b_hi + a_hi*b_lo + a_lo*b_hi
// (the a_lo*b_lo term is below 2^-16 relative), accumulated in FP32.
// ============================================================================

typedef __attribute__((ext_vector_type(2))) float f32x2_t;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2_t;

// 8 consecutive bf16 of a column starting at off (rows r0..r0+7 masked against rend).
__device__ __forceinline__ u32x4_t load_raw8(const uint16_t* __restrict__ p, int64_t off, int64_t r0, int64_t rend,
                                             bool colok, bool vec) {
  if (colok && vec && r0 + 8 <= rend) return __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p + off));
  u32x4_t v = {0u, 0u, 0u, 0u};
  if (colok) {
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (r0 + e < rend) v[e / 2] |= uint32_t(p[off + e]) << (16 * (e % 2));
  }
  return v;
}

// hi/lo bf16 split of the FP32 differences of two packed bf16x8 vectors.
__device__ __forceinline__ void diff_split8(u32x4_t x, u32x4_t y, u32x4_t& hi, u32x4_t& lo) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    f32x2_t d;
    d[0] = __uint_as_float(x[q] << 16) - __uint_as_float(y[q] << 16);
    d[1] = __uint_as_float(x[q] & 0xFFFF0000u) - __uint_as_float(y[q] & 0xFFFF0000u);
    const bf16x2_t h = __builtin_convertvector(d, bf16x2_t);
    const f32x2_t hf = __builtin_convertvector(h, f32x2_t);
    const f32x2_t r = d - hf;
    const bf16x2_t l = __builtin_convertvector(r, bf16x2_t);
    hi[q] = __builtin_bit_cast(uint32_t, h);
    lo[q] = __builtin_bit_cast(uint32_t, l);
  }
}

// pass T: tile = 32 columns of dA per workgroup; an iteration covers 64 rows.
// Loads are fully coalesced (8 lanes read the 128 contiguous bytes of one
// column, a wave 8 columns per instruction); the hi/lo differences are staged
// column-major in the wave's LDS region (column pitch 64 + 8 bf16 keeps the
// 16-byte fragment reads conflict-free) and read back as lane (column lc,
// half h): rows 32h + 8q .. 32h + 8q + 7 for k-step q.
// B panel (U) in fragment order: for rows block b (64 rows), k-step q, half h,
// column c: 8 consecutive bf16 at ((((b*4 + q)*2 + h)*rb + c)*8).
constexpr int kTPitch = 72;

template <int CB>
__global__ __launch_bounds__(kPassThreads) void k_passT_bf16x3(const uint16_t* __restrict__ an,
                                                               const uint16_t* __restrict__ ap, int64_t lda,
                                                               int64_t n, int64_t m, const uint16_t* __restrict__ Bhi,
                                                               const uint16_t* __restrict__ Blo, int64_t ldb,
                                                               int rb, int64_t ichunk, float* __restrict__ part,
                                                               int vec) {
  constexpr int NV = CB * 16;
  constexpr int STAGE = 4 * 2 * 32 * kTPitch * 2;  // bytes
  constexpr int RED = 3 * NV * 64 * 4;
  __shared__ __attribute__((aligned(16))) unsigned char smem[STAGE > RED ? STAGE : RED];
  (void)ldb;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 31, h = lane >> 5;
  const int64_t j0 = (int64_t)blockIdx.x * 32;
  const int s = blockIdx.y;
  const int64_t ib = (int64_t)s * ichunk;
  const int64_t ie = min(n, ib + ichunk);
  const WaveRange R = wave_range(ib, ie, wave, 64);
  const bool v = vec != 0;
  uint16_t* th = reinterpret_cast<uint16_t*>(smem) + (wave * 2 + 0) * 32 * kTPitch;
  uint16_t* tl = reinterpret_cast<uint16_t*>(smem) + (wave * 2 + 1) * 32 * kTPitch;
  f32x16_t acc[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) acc[c] = f32x16_t{};
  // load lane mapping: instruction p covers columns 8p .. 8p+7; lane -> column 8p + lane/8, rows 8*(lane%8)
  const int lcol = lane >> 3, loct = lane & 7;
  u32x4_t xn[4], xp[4], nxn[4], nxp[4];
  auto load = [&](int64_t i, u32x4_t (&A1)[4], u32x4_t (&A0)[4]) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int64_t jj = j0 + 8 * p + lcol;
      const int64_t r0 = i + 8 * loct;
      A1[p] = load_raw8(an, jj * lda + r0, r0, ie, jj < m, v);
      A0[p] = load_raw8(ap, jj * lda + r0, r0, ie, jj < m && ap != nullptr, v);  // no ap: A itself
    }
  };
  if (R.b < R.e) load(R.b, xn, xp);
  for (int64_t i = R.b; i < R.e; i += 64) {
    const bool more = i + 64 < R.e;
    if (more) load(i + 64, nxn, nxp);
    // B fragments of this 64-row block (L2-resident panel)
    u32x4_t bh[4][CB], bl[4][CB];
    const int64_t blk = i >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        const int64_t off = ((((blk * 4 + q) * 2 + h) * rb) + c * 32 + lc) * 8;
        bh[q][c] = *reinterpret_cast<const u32x4_t*>(Bhi + off);
        bl[q][c] = *reinterpret_cast<const u32x4_t*>(Blo + off);
      }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      u32x4_t hv, lv;
      diff_split8(xn[p], xp[p], hv, lv);
      const int col = 8 * p + lcol;
      *reinterpret_cast<u32x4_t*>(th + col * kTPitch + 8 * loct) = hv;
      *reinterpret_cast<u32x4_t*>(tl + col * kTPitch + 8 * loct) = lv;
    }
    wave_lds_fence();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const u32x4_t ah = *reinterpret_cast<const u32x4_t*>(th + lc * kTPitch + 32 * h + 8 * q);
      const u32x4_t al = *reinterpret_cast<const u32x4_t*>(tl + lc * kTPitch + 32 * h + 8 * q);
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        acc[c] = mfma_bf16(as_bf16x8(ah), as_bf16x8(bh[q][c]), acc[c]);
        acc[c] = mfma_bf16(as_bf16x8(ah), as_bf16x8(bl[q][c]), acc[c]);
        acc[c] = mfma_bf16(as_bf16x8(al), as_bf16x8(bh[q][c]), acc[c]);
      }
    }
    wave_lds_fence();
    if (more) {
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        xn[p] = nxn[p];
        xp[p] = nxp[p];
      }
    }
  }
  float flat[NV];
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int w = 0; w < 16; ++w) flat[c * 16 + w] = acc[c][w];
  wg_sum<float, NV>(flat, reinterpret_cast<float*>(smem), wave, lane);
  if (wave != 0) return;
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      const int64_t jj = j0 + 8 * (w / 4) + 4 * h + (w % 4);
      if (jj < m) part[((int64_t)s * rb + c * 32 + lc) * m + jj] = flat[c * 16 + w];
    }
}

// pass T, wide form (rb = 32): a workgroup covers 128 columns of dA, wave w the
// columns 32w .. 32w+31, all four over the same rows.  Each 64-row block of the
// B panel (hi and lo, 8 KB, contiguous in the panel) is then loaded once per
// workgroup into LDS (double-buffered, prefetched into registers one block
// ahead) and read by the four waves, instead of once per wave from L2.  The A
// tiles are loaded, split and staged per wave as in k_passT_bf16x3; kSingle
// streams A[k_next] itself (two products per k-step, see k_passT_bf16_single).
// No cross-wave reduction: every wave owns its 32 output rows.
template <bool kSingle>
__global__ __launch_bounds__(kPassThreads) void k_passT_bf16_wide(const uint16_t* __restrict__ an,
                                                                  const uint16_t* __restrict__ ap, int64_t lda,
                                                                  int64_t n, int64_t m,
                                                                  const uint16_t* __restrict__ Bhi,
                                                                  const uint16_t* __restrict__ Blo, int rb,
                                                                  int64_t ichunk, float* __restrict__ part,
                                                                  int vec) {
  __shared__ __attribute__((aligned(16))) uint16_t sA[4][2][32 * kTPitch];
  __shared__ u32x4_t sB[2][2][256];  // [buffer][hi, lo][64 rows x 32 columns in fragment order]
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int lc = lane & 31, h = lane >> 5;
  const int64_t j0 = (int64_t)blockIdx.x * 128 + 32 * wave;
  const int s = blockIdx.y;
  const int64_t ib = (int64_t)s * ichunk;
  const int64_t ie = min(n, ib + ichunk);
  const bool v = vec != 0;
  uint16_t* th = sA[wave][0];
  uint16_t* tl = sA[wave][1];
  f32x16_t acc = f32x16_t{};
  const int lcol = lane >> 3, loct = lane & 7;
  u32x4_t xn[4], xp[4], nxn[4], nxp[4];
  auto load = [&](int64_t i, u32x4_t (&A1)[4], u32x4_t (&A0)[4]) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int64_t jj = j0 + 8 * p + lcol;
      const int64_t r0 = i + 8 * loct;
      A1[p] = load_raw8(an, jj * lda + r0, r0, ie, jj < m, v);
      if (!kSingle) A0[p] = load_raw8(ap, jj * lda + r0, r0, ie, jj < m, v);
    }
  };
  // the B block of rows i .. i + 63: 256 hi and 256 lo 16-byte pieces, one of each per thread
  auto loadB = [&](int64_t i, u32x4_t& gh, u32x4_t& gl) {
    const int64_t base = (i >> 6) * 64 * rb;
    gh = reinterpret_cast<const u32x4_t*>(Bhi + base)[t];
    gl = reinterpret_cast<const u32x4_t*>(Blo + base)[t];
  };
  u32x4_t gh, gl;
  int buf = 0;
  if (ib < ie) {
    load(ib, xn, xp);
    loadB(ib, gh, gl);
    sB[0][0][t] = gh;
    sB[0][1][t] = gl;
  }
  __syncthreads();
  for (int64_t i = ib; i < ie; i += 64) {
    const bool more = i + 64 < ie;
    if (more) {
      load(i + 64, nxn, nxp);
      loadB(i + 64, gh, gl);
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int col = 8 * p + lcol;
      if (kSingle) {
        *reinterpret_cast<u32x4_t*>(th + col * kTPitch + 8 * loct) = xn[p];
      } else {
        u32x4_t hv, lv;
        diff_split8(xn[p], xp[p], hv, lv);
        *reinterpret_cast<u32x4_t*>(th + col * kTPitch + 8 * loct) = hv;
        *reinterpret_cast<u32x4_t*>(tl + col * kTPitch + 8 * loct) = lv;
      }
    }
    wave_lds_fence();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const u32x4_t ah = *reinterpret_cast<const u32x4_t*>(th + lc * kTPitch + 32 * h + 8 * q);
      const u32x4_t bh = sB[buf][0][(q * 2 + h) * 32 + lc];
      const u32x4_t bl = sB[buf][1][(q * 2 + h) * 32 + lc];
      acc = mfma_bf16(as_bf16x8(ah), as_bf16x8(bh), acc);
      acc = mfma_bf16(as_bf16x8(ah), as_bf16x8(bl), acc);
      if (!kSingle) {
        const u32x4_t al = *reinterpret_cast<const u32x4_t*>(tl + lc * kTPitch + 32 * h + 8 * q);
        acc = mfma_bf16(as_bf16x8(al), as_bf16x8(bh), acc);
      }
    }
    if (more) {
      sB[buf ^ 1][0][t] = gh;
      sB[buf ^ 1][1][t] = gl;
    }
    __syncthreads();  // sB[buf] read by all waves, sB[buf ^ 1] written; the A staging is free again
    if (more) {
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        xn[p] = nxn[p];
        if (!kSingle) xp[p] = nxp[p];
      }
    }
    buf ^= 1;
  }
#pragma unroll
  for (int w = 0; w < 16; ++w) {
    const int64_t jj = j0 + 8 * (w / 4) + 4 * h + (w % 4);
    if (jj < m) part[((int64_t)s * rb + lc) * m + jj] = acc[w];
  }
}

// pass N: the operand needs 8 consecutive columns per lane but memory runs
// along rows, so each wave stages a (32*TM rows x 16 columns) hi/lo tile in
// its own LDS region: coalesced 16-byte loads down the columns, stored
// column-major, read back as 8 column-consecutive bf16 per lane.
// B panel (V) in fragment order: for the 16-row block b, half h, column c:
// 8 consecutive bf16 at (((b*2 + h)*rb + c)*8).
constexpr int kNPad = 4;  // LDS column stride = rows + 4 (bf16): the two lane halves hit different banks

template <int TM, int CB>
struct PassNLds {
  static constexpr int LDC = 32 * TM + kNPad;
  static constexpr int STAGE_BYTES = 4 * 2 * 16 * LDC * 2;
  static constexpr int RED_BYTES = 3 * TM * CB * 16 * 64 * 4;
  static constexpr int BYTES = STAGE_BYTES > RED_BYTES ? STAGE_BYTES : RED_BYTES;
};

template <int TM, int CB>
__global__ __launch_bounds__(kPassThreads) void k_passN_bf16x3(const uint16_t* __restrict__ an,
                                                               const uint16_t* __restrict__ ap, int64_t lda,
                                                               int64_t n, int64_t m, const uint16_t* __restrict__ Bhi,
                                                               const uint16_t* __restrict__ Blo, int64_t ldb,
                                                               int rb, int64_t jchunk, float* __restrict__ part,
                                                               int vec) {
  using L = PassNLds<TM, CB>;
  constexpr int LDC = L::LDC;
  constexpr int NV = TM * CB * 16;
  __shared__ __attribute__((aligned(16))) unsigned char smem[L::BYTES];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 31, h = lane >> 5;
  const int64_t i0 = (int64_t)blockIdx.x * (32 * TM);
  const int s = blockIdx.y;
  const int64_t jb = (int64_t)s * jchunk;
  const int64_t je = min(m, jb + jchunk);
  const WaveRange R = wave_range(jb, je, wave, 16);
  const bool v = vec != 0;
  uint16_t* th = reinterpret_cast<uint16_t*>(smem) + (wave * 2 + 0) * 16 * LDC;
  uint16_t* tl = reinterpret_cast<uint16_t*>(smem) + (wave * 2 + 1) * 16 * LDC;
  f32x16_t acc[TM][CB];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[t][c] = f32x16_t{};
  // staging chunk q of a k-step: b = q*64 + lane -> column b / (4*TM), row octet b % (4*TM)
  u32x4_t xn[TM], xp[TM], bh[CB], bl[CB], nxn[TM], nxp[TM], nbh[CB], nbl[CB];
  auto load = [&](int64_t j, u32x4_t (&A1)[TM], u32x4_t (&A0)[TM], u32x4_t (&H)[CB], u32x4_t (&Lo)[CB]) {
#pragma unroll
    for (int q = 0; q < TM; ++q) {
      const int b = q * 64 + lane;
      const int col = b / (4 * TM), oct = b % (4 * TM);
      const int64_t jj = j + col;
      const int64_t r0 = i0 + 8 * oct;
      A1[q] = load_raw8(an, jj * lda + r0, r0, n, jj < je, v);
      A0[q] = load_raw8(ap, jj * lda + r0, r0, n, jj < je && ap != nullptr, v);  // no ap: A itself
    }
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int64_t boff = ((((j >> 4) * 2 + h) * rb) + c * 32 + lc) * 8;
      H[c] = *reinterpret_cast<const u32x4_t*>(Bhi + boff);
      Lo[c] = *reinterpret_cast<const u32x4_t*>(Blo + boff);
    }
  };
  if (R.b < R.e) load(R.b, xn, xp, bh, bl);
  for (int64_t j = R.b; j < R.e; j += 16) {
    const bool more = j + 16 < R.e;
    if (more) load(j + 16, nxn, nxp, nbh, nbl);
#pragma unroll
    for (int q = 0; q < TM; ++q) {
      const int b = q * 64 + lane;
      const int col = b / (4 * TM), oct = b % (4 * TM);
      u32x4_t hv, lv;
      diff_split8(xn[q], xp[q], hv, lv);
      // LDC*2 bytes is 8-byte aligned: two 8-byte stores per 16-byte chunk
      uint2* ph = reinterpret_cast<uint2*>(th + col * LDC + 8 * oct);
      uint2* pl = reinterpret_cast<uint2*>(tl + col * LDC + 8 * oct);
      ph[0] = uint2{hv[0], hv[1]};
      ph[1] = uint2{hv[2], hv[3]};
      pl[0] = uint2{lv[0], lv[1]};
      pl[1] = uint2{lv[2], lv[3]};
    }
    wave_lds_fence();
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
    wave_lds_fence();
    if (more) {
#pragma unroll
      for (int q = 0; q < TM; ++q) {
        xn[q] = nxn[q];
        xp[q] = nxp[q];
      }
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        bh[c] = nbh[c];
        bl[c] = nbl[c];
      }
    }
  }
  float flat[NV];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int w = 0; w < 16; ++w) flat[(t * CB + c) * 16 + w] = acc[t][c][w];
  wg_sum<float, NV>(flat, reinterpret_cast<float*>(smem), wave, lane);
  if (wave != 0) return;
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int w = 0; w < 16; ++w) {
        const int64_t ii = i0 + 32 * t + 8 * (w / 4) + 4 * h + (w % 4);
        if (ii < n) part[((int64_t)s * rb + c * 32 + lc) * n + ii] = flat[(t * CB + c) * 16 + w];
      }
}

// ----------------------------------------------------------------------------
// Single-operand forms (A itself, no previous snapshot): A is exact in BF16, so
// its hi part is the raw value and its lo part vanishes (2 products per k-step
// instead of 3), and the load slots the difference would use fetch the next
// tile instead, keeping as many bytes in flight per wave as the two-operand
// passes.  Pass N takes 32 columns per iteration, pass T 128 rows.
// ----------------------------------------------------------------------------

template <int TM, int CB>
__global__ __launch_bounds__(kPassThreads) void k_passN_bf16_single(const uint16_t* __restrict__ an, int64_t lda,
                                                                    int64_t n, int64_t m,
                                                                    const uint16_t* __restrict__ Bhi,
                                                                    const uint16_t* __restrict__ Blo, int rb,
                                                                    int64_t jchunk, float* __restrict__ part,
                                                                    int vec) {
  using L = PassNLds<TM, CB>;
  constexpr int LDC = L::LDC;
  constexpr int NV = TM * CB * 16;
  __shared__ __attribute__((aligned(16))) unsigned char smem[L::BYTES];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 31, h = lane >> 5;
  const int64_t i0 = (int64_t)blockIdx.x * (32 * TM);
  const int s = blockIdx.y;
  const int64_t jb = (int64_t)s * jchunk;
  const int64_t je = min(m, jb + jchunk);
  const WaveRange R = wave_range(jb, je, wave, 32);
  const bool v = vec != 0;
  uint16_t* th = reinterpret_cast<uint16_t*>(smem) + (wave * 2 + 0) * 16 * LDC;
  f32x16_t acc[TM][CB];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[t][c] = f32x16_t{};
  // tile g (0: columns j..j+15, 1: j+16..j+31) with its B fragments
  u32x4_t xa[2][TM], nxa[2][TM], bh[2][CB], bl[2][CB], nbh[2][CB], nbl[2][CB];
  auto load = [&](int64_t j, u32x4_t (&X)[2][TM], u32x4_t (&H)[2][CB], u32x4_t (&Lo)[2][CB]) {
#pragma unroll
    for (int g = 0; g < 2; ++g) {
#pragma unroll
      for (int q = 0; q < TM; ++q) {
        const int b = q * 64 + lane;
        const int col = b / (4 * TM), oct = b % (4 * TM);
        const int64_t jj = j + 16 * g + col;
        const int64_t r0 = i0 + 8 * oct;
        X[g][q] = load_raw8(an, jj * lda + r0, r0, n, jj < je, v);
      }
      // the second tile's fragments exist only inside the wave's range (the panel ends there)
      const bool bok = j + 16 * g < R.e;
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        const int64_t boff = (((((j >> 4) + g) * 2 + h) * rb) + c * 32 + lc) * 8;
        H[g][c] = bok ? *reinterpret_cast<const u32x4_t*>(Bhi + boff) : u32x4_t{0u, 0u, 0u, 0u};
        Lo[g][c] = bok ? *reinterpret_cast<const u32x4_t*>(Blo + boff) : u32x4_t{0u, 0u, 0u, 0u};
      }
    }
  };
  if (R.b < R.e) load(R.b, xa, bh, bl);
  for (int64_t j = R.b; j < R.e; j += 32) {
    const bool more = j + 32 < R.e;
    if (more) load(j + 32, nxa, nbh, nbl);
#pragma unroll
    for (int g = 0; g < 2; ++g) {
#pragma unroll
      for (int q = 0; q < TM; ++q) {
        const int b = q * 64 + lane;
        const int col = b / (4 * TM), oct = b % (4 * TM);
        uint2* ph = reinterpret_cast<uint2*>(th + col * LDC + 8 * oct);
        ph[0] = uint2{xa[g][q][0], xa[g][q][1]};
        ph[1] = uint2{xa[g][q][2], xa[g][q][3]};
      }
      wave_lds_fence();
#pragma unroll
      for (int t = 0; t < TM; ++t) {
        const int row = 32 * t + lc;
        u32x4_t ah;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int c0 = 8 * h + 2 * q;
          ah[q] = uint32_t(th[c0 * LDC + row]) | (uint32_t(th[(c0 + 1) * LDC + row]) << 16);
        }
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          acc[t][c] = mfma_bf16(as_bf16x8(ah), as_bf16x8(bh[g][c]), acc[t][c]);
          acc[t][c] = mfma_bf16(as_bf16x8(ah), as_bf16x8(bl[g][c]), acc[t][c]);
        }
      }
      wave_lds_fence();
    }
    if (more) {
#pragma unroll
      for (int g = 0; g < 2; ++g) {
#pragma unroll
        for (int q = 0; q < TM; ++q) xa[g][q] = nxa[g][q];
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          bh[g][c] = nbh[g][c];
          bl[g][c] = nbl[g][c];
        }
      }
    }
  }
  float flat[NV];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int w = 0; w < 16; ++w) flat[(t * CB + c) * 16 + w] = acc[t][c][w];
  wg_sum<float, NV>(flat, reinterpret_cast<float*>(smem), wave, lane);
  if (wave != 0) return;
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int w = 0; w < 16; ++w) {
        const int64_t ii = i0 + 32 * t + 8 * (w / 4) + 4 * h + (w % 4);
        if (ii < n) part[((int64_t)s * rb + c * 32 + lc) * n + ii] = flat[(t * CB + c) * 16 + w];
      }
}

template <int CB>
__global__ __launch_bounds__(kPassThreads) void k_passT_bf16_single(const uint16_t* __restrict__ an, int64_t lda,
                                                                    int64_t n, int64_t m,
                                                                    const uint16_t* __restrict__ Bhi,
                                                                    const uint16_t* __restrict__ Blo, int rb,
                                                                    int64_t ichunk, float* __restrict__ part,
                                                                    int vec) {
  constexpr int NV = CB * 16;
  constexpr int STAGE = 4 * 2 * 32 * kTPitch * 2;  // bytes
  constexpr int RED = 3 * NV * 64 * 4;
  __shared__ __attribute__((aligned(16))) unsigned char smem[STAGE > RED ? STAGE : RED];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 31, h = lane >> 5;
  const int64_t j0 = (int64_t)blockIdx.x * 32;
  const int s = blockIdx.y;
  const int64_t ib = (int64_t)s * ichunk;
  const int64_t ie = min(n, ib + ichunk);
  const WaveRange R = wave_range(ib, ie, wave, 128);
  const bool v = vec != 0;
  uint16_t* th = reinterpret_cast<uint16_t*>(smem) + (wave * 2 + 0) * 32 * kTPitch;
  f32x16_t acc[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) acc[c] = f32x16_t{};
  const int lcol = lane >> 3, loct = lane & 7;
  // tile g: rows i + 64 g .. i + 64 g + 63, with the B fragments of its 64-row block
  // (loaded only inside the wave's range: the panel ends there)
  u32x4_t xa[2][4], nxa[2][4], bh[2][4][CB], bl[2][4][CB], nbh[2][4][CB], nbl[2][4][CB];
  auto load = [&](int64_t i, u32x4_t (&X)[2][4], u32x4_t (&H)[2][4][CB], u32x4_t (&Lo)[2][4][CB]) {
#pragma unroll
    for (int g = 0; g < 2; ++g) {
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int64_t jj = j0 + 8 * p + lcol;
        const int64_t r0 = i + 64 * g + 8 * loct;
        X[g][p] = load_raw8(an, jj * lda + r0, r0, ie, jj < m, v);
      }
      const bool bok = i + 64 * g < R.e;
      const int64_t blk = (i >> 6) + g;
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          const int64_t off = ((((blk * 4 + q) * 2 + h) * rb) + c * 32 + lc) * 8;
          H[g][q][c] = bok ? *reinterpret_cast<const u32x4_t*>(Bhi + off) : u32x4_t{0u, 0u, 0u, 0u};
          Lo[g][q][c] = bok ? *reinterpret_cast<const u32x4_t*>(Blo + off) : u32x4_t{0u, 0u, 0u, 0u};
        }
    }
  };
  if (R.b < R.e) load(R.b, xa, bh, bl);
  for (int64_t i = R.b; i < R.e; i += 128) {
    const bool more = i + 128 < R.e;
    if (more) load(i + 128, nxa, nbh, nbl);
#pragma unroll
    for (int g = 0; g < 2; ++g) {
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int col = 8 * p + lcol;
        *reinterpret_cast<u32x4_t*>(th + col * kTPitch + 8 * loct) = xa[g][p];
      }
      wave_lds_fence();
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const u32x4_t ah = *reinterpret_cast<const u32x4_t*>(th + lc * kTPitch + 32 * h + 8 * q);
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          acc[c] = mfma_bf16(as_bf16x8(ah), as_bf16x8(bh[g][q][c]), acc[c]);
          acc[c] = mfma_bf16(as_bf16x8(ah), as_bf16x8(bl[g][q][c]), acc[c]);
        }
      }
      wave_lds_fence();
    }
    if (more) {
#pragma unroll
      for (int g = 0; g < 2; ++g) {
#pragma unroll
        for (int p = 0; p < 4; ++p) xa[g][p] = nxa[g][p];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int c = 0; c < CB; ++c) {
            bh[g][q][c] = nbh[g][q][c];
            bl[g][q][c] = nbl[g][q][c];
          }
      }
    }
  }
  float flat[NV];
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int w = 0; w < 16; ++w) flat[c * 16 + w] = acc[c][w];
  wg_sum<float, NV>(flat, reinterpret_cast<float*>(smem), wave, lane);
  if (wave != 0) return;
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      const int64_t jj = j0 + 8 * (w / 4) + 4 * h + (w % 4);
      if (jj < m) part[((int64_t)s * rb + c * 32 + lc) * m + jj] = flat[c * 16 + w];
    }
}

// ============================================================================
// planning and launch
// ============================================================================

// Tile height (output rows per workgroup) in units of the instruction's M
// (16 for FP64, 32 otherwise); accumulators are kept at <= 8 (FP64) or 4
// (FP32) blocks per wave.
static int tile_units(int mode, bool transposed, int rb) {
  if (mode == PASS_F64) {
    int cb = rb / 16;
    return cb <= 2 ? 4 : (cb <= 4 ? 2 : 1);
  }
  int cb = rb / 32;
  if (mode == PASS_F32) return transposed ? 4 : (cb <= 1 ? 4 : (cb == 2 ? 2 : 1));  // pass T: the wide form
  if (transposed) return cb <= 1 ? 4 : 1;  // rb = 32: the wide pass T (128 columns per workgroup)
  // rb = 32: 64-row tiles (128 VGPRs + 32 AGPRs, 24 KB of LDS: four waves per SIMD) keep
  // twice the loads in flight per CU of the 128-row tiles (two waves per SIMD); at
  // n = m = 16384 pass N took 175 us against 192-204 us
  return cb <= 1 ? 2 : (cb == 2 ? 2 : 1);
}

static int wg_span(int mode, bool transposed, int rb) {
  return (mode == PASS_F64 ? 16 : 32) * tile_units(mode, transposed, rb);
}

constexpr int64_t kChunkQuantum = 256;  // 4 waves x the largest per-wave iteration (64 rows)

// The reduction chunk for about `target` workgroups, with at least 1024 of the
// reduction range per split.
static int64_t chunk_for_target(int mode, bool transposed, int64_t out_rows, int64_t red_len, int rb,
                                int64_t target) {
  const int span = wg_span(mode, transposed, rb);
  const int64_t tiles = (out_rows + span - 1) / span;
  int64_t want = (target + tiles - 1) / tiles;
  const int64_t maxs = (red_len + 1023) / 1024;
  if (want > maxs) want = maxs;
  if (want < 1) want = 1;
  const int64_t ch = (red_len + want - 1) / want;
  return (ch + kChunkQuantum - 1) / kChunkQuantum * kChunkQuantum;
}

int64_t plan_pass_chunk(int64_t out_rows, int64_t red_len, int rb, int64_t ch, int* splits, int64_t* chunk) {
  ch = (ch + kChunkQuantum - 1) / kChunkQuantum * kChunkQuantum;
  if (ch < kChunkQuantum) ch = kChunkQuantum;
  const int64_t s = (red_len + ch - 1) / ch;
  *splits = (int)s;
  *chunk = ch;
  return s * rb * out_rows;
}

// the default: ~512 workgroups (2048 waves); ~768 for the BF16 pass N at rb <= 32,
// whose 64-row workgroups run four waves per SIMD (at n = m = 16384: 3 splits, 170
// us against 196 us for 2)
int64_t plan_pass(int mode, bool transposed, int64_t out_rows, int64_t red_len, int rb, int* splits,
                  int64_t* chunk) {
  const int64_t target = (mode == PASS_BF16X3 && !transposed && rb <= 32) ? 768 : 512;
  return plan_pass_chunk(out_rows, red_len, rb, chunk_for_target(mode, transposed, out_rows, red_len, rb, target),
                         splits, chunk);
}

int pass_chunk_candidates(int mode, bool transposed, int64_t out_rows, int64_t red_len, int rb, int64_t* chunks,
                          int cap) {
  int k = 0;
  for (int64_t target : {128, 192, 256, 384, 512, 640, 768, 1024, 1280, 1536, 2048, 3072}) {
    const int64_t ch = chunk_for_target(mode, transposed, out_rows, red_len, rb, target);
    bool dup = false;
    for (int i = 0; i < k; ++i) dup = dup || chunks[i] == ch;
    if (!dup && k < cap) chunks[k++] = ch;
  }
  return k;
}

// vector loads of 4 consecutive elements need 4-element-aligned columns
static int vec_f64(const PassArgs& a, size_t esz) {
  const size_t al = esz == 8 ? 16 : 4 * esz;
  return (a.lda % 4 == 0) && ((uintptr_t)a.anext % al == 0) && ((uintptr_t)a.aprev % al == 0) ? 1 : 0;
}

// vector loads of TM consecutive floats need TM-aligned columns
static int vec_f32(const PassArgs& a, int tm) {
  if (tm <= 1) return 0;
  const size_t al = 4 * (size_t)tm;
  return (a.lda % tm == 0) && ((uintptr_t)a.anext % al == 0) && ((uintptr_t)a.aprev % al == 0) ? 1 : 0;
}

static bool vec_ok(const PassArgs& a) {
  return (a.lda % 8 == 0) && ((uintptr_t)a.anext % 16 == 0) && ((uintptr_t)a.aprev % 16 == 0);
}

template <int TM, int CB>
static hipError_t run_pass(int mode, bool transposed, const PassArgs& a, hipStream_t st) {
  const int64_t out_rows = transposed ? a.m : a.n;
  const int span = (mode == PASS_F64 ? 16 : 32) * TM;
  dim3 grid((unsigned)((out_rows + span - 1) / span), (unsigned)a.splits);
  dim3 blk(kPassThreads);
  if (mode == PASS_F64) {
    const double* B = (const double*)a.B;
    double* P = (double*)a.part;
#define LRI_F64_LAUNCH(T)                                                                                    \
  if (transposed)                                                                                            \
    hipLaunchKernelGGL((k_passT_f64<TM, CB, T>), grid, blk, 0, st, (const T*)a.anext, (const T*)a.aprev,      \
                       a.lda, a.n, a.m, B, a.rb, a.chunk, P, vec_f64(a, sizeof(T)));                         \
  else                                                                                                       \
    hipLaunchKernelGGL((k_passN_f64<TM, CB, T>), grid, blk, 0, st, (const T*)a.anext, (const T*)a.aprev, a.lda,\
                       a.n, a.m, B, a.rb, a.chunk, P, vec_f64(a, sizeof(T)));
    if (a.dtype == LRI_F64) {
      LRI_F64_LAUNCH(double)
    } else if (a.dtype == LRI_F32) {
      LRI_F64_LAUNCH(float)
    } else {
      LRI_F64_LAUNCH(uint16_t)
    }
#undef LRI_F64_LAUNCH
  } else if (mode == PASS_F32) {
    const float* an = (const float*)a.anext;
    const float* ap = (const float*)a.aprev;
    if (transposed) {
      // tile = 128 columns whatever TM (the wide form)
      const dim3 gw((unsigned)((a.m + 127) / 128), (unsigned)a.splits);
      const int vw = vec_f32(a, 4);
      hipLaunchKernelGGL((k_passT_f32_wide<CB>), gw, blk, 0, st, an, ap, a.lda, a.n, a.m, (const float*)a.B, a.rb,
                         a.chunk, (float*)a.part, vw);
    } else {
      hipLaunchKernelGGL((k_passN_f32<TM, CB>), grid, blk, 0, st, an, ap, a.lda, a.n, a.m, (const float*)a.B, a.rb,
                         a.chunk, (float*)a.part, vec_f32(a, TM));
    }
  } else {
    const uint16_t* an = (const uint16_t*)a.anext;
    const uint16_t* ap = (const uint16_t*)a.aprev;
    const int v = vec_ok(a) ? 1 : 0;
    if constexpr (TM == 4 && CB == 1) {
      if (transposed) {  // rb = 32: the wide pass T
        if (!ap)
          hipLaunchKernelGGL((k_passT_bf16_wide<true>), grid, blk, 0, st, an, an, a.lda, a.n, a.m,
                             (const uint16_t*)a.B, (const uint16_t*)a.B2, a.rb, a.chunk, (float*)a.part, v);
        else
          hipLaunchKernelGGL((k_passT_bf16_wide<false>), grid, blk, 0, st, an, ap, a.lda, a.n, a.m,
                             (const uint16_t*)a.B, (const uint16_t*)a.B2, a.rb, a.chunk, (float*)a.part, v);
        return hipGetLastError();
      }
    }
    if (!ap) {
      if (transposed)
        hipLaunchKernelGGL((k_passT_bf16_single<CB>), grid, blk, 0, st, an, a.lda, a.n, a.m, (const uint16_t*)a.B,
                           (const uint16_t*)a.B2, a.rb, a.chunk, (float*)a.part, v);
      else
        hipLaunchKernelGGL((k_passN_bf16_single<TM, CB>), grid, blk, 0, st, an, a.lda, a.n, a.m,
                           (const uint16_t*)a.B, (const uint16_t*)a.B2, a.rb, a.chunk, (float*)a.part, v);
      return hipGetLastError();
    }
    if (transposed)
      hipLaunchKernelGGL((k_passT_bf16x3<CB>), grid, blk, 0, st, an, ap, a.lda, a.n, a.m,
                         (const uint16_t*)a.B, (const uint16_t*)a.B2, a.ldb, a.rb, a.chunk, (float*)a.part, v);
    else
      hipLaunchKernelGGL((k_passN_bf16x3<TM, CB>), grid, blk, 0, st, an, ap, a.lda, a.n, a.m,
                         (const uint16_t*)a.B, (const uint16_t*)a.B2, a.ldb, a.rb, a.chunk, (float*)a.part, v);
  }
  return hipGetLastError();
}

static hipError_t launch_pass(int mode, bool transposed, const PassArgs& a, hipStream_t st) {
  if (current_recorder()) return hipErrorNotSupported;  // the passes have no batched form
  if (mode == PASS_BF16X3 && a.dtype != LRI_BF16) return hipErrorInvalidValue;
  if (mode == PASS_F32 && a.dtype != LRI_F32) return hipErrorInvalidValue;
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
    case 32:
      return tu == 4 ? run_pass<4, 1>(mode, transposed, a, st)
                     : (tu == 2 ? run_pass<2, 1>(mode, transposed, a, st) : run_pass<1, 1>(mode, transposed, a, st));
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

// ============================================================================
// Streaming-read probe: what bare non-temporal 16-byte loads reach over two
// column-major 16384 x 16384 BF16 matrices (no arithmetic), as context for the
// passes' fraction of the HBM peak.  pattern 0: contiguous; 1: a wave
// instruction covers 8 columns x 128 B (pass T's pattern); 2: 1 column x 1 KB.
// 512 workgroups of 256 threads, each over 32 columns (pattern 0: a grid-stride
// sweep of the same bytes).
// ============================================================================

constexpr int64_t kProbeN = 16384;

template <int CPW>
__global__ __launch_bounds__(256) void k_stream_probe(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b,
                                                      uint32_t* __restrict__ sink) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t acc = 0;
  if (CPW == 0) {
    const int64_t n16 = kProbeN * kProbeN / 8;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256) {
      const u32x4_t x = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(a) + i);
      const u32x4_t y = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(b) + i);
      acc ^= x[0] ^ x[1] ^ x[2] ^ x[3] ^ y[0] ^ y[1] ^ y[2] ^ y[3];
    }
  } else {
    constexpr int lpc = 64 / (CPW > 0 ? CPW : 1);  // lanes per column
    constexpr int rpi = lpc * 8;                     // rows per column per instruction
    const int64_t c0 = (int64_t)blockIdx.x * 32;
    for (int cg = wave * CPW; cg < 32; cg += 4 * CPW) {
      const int64_t col = c0 + cg + lane / lpc;
      const int64_t rbase = (lane % lpc) * 8;
      for (int64_t r = 0; r < kProbeN; r += 4 * rpi) {
        u32x4_t x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t off = col * kProbeN + r + (int64_t)u * rpi + rbase;
          x[u] = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(a + off));
          y[u] = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(b + off));
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc ^= x[u][0] ^ x[u][1] ^ x[u][2] ^ x[u][3] ^ y[u][0] ^ y[u][1] ^ y[u][2] ^ y[u][3];
      }
    }
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the loads
}

hipError_t stream_probe(int pattern, double* tbps) {
  if (pattern < 0 || pattern > 2 || !tbps) return hipErrorInvalidValue;
  const size_t bytes = (size_t)kProbeN * kProbeN * 2;
  uint16_t *a = nullptr, *b = nullptr;
  uint32_t* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&a, bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&b, bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&sink, 64);
  if (e == hipSuccess) e = hipMemset(a, 1, bytes);
  if (e == hipSuccess) e = hipMemset(b, 2, bytes);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  float ms = 0.f;
  const dim3 grid((unsigned)(kProbeN / 32)), blk(256);
  auto launch = [&] {
    if (pattern == 0) hipLaunchKernelGGL(k_stream_probe<0>, grid, blk, 0, 0, a, b, sink);
    else if (pattern == 1) hipLaunchKernelGGL(k_stream_probe<8>, grid, blk, 0, 0, a, b, sink);
    else hipLaunchKernelGGL(k_stream_probe<1>, grid, blk, 0, 0, a, b, sink);
    return hipGetLastError();
  };
  for (int it = 0; it < 3 && e == hipSuccess; ++it) e = launch();
  if (e == hipSuccess) e = hipEventRecord(e0, 0);
  for (int it = 0; it < 10 && e == hipSuccess; ++it) e = launch();
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) *tbps = 2.0 * bytes / (ms / 10 * 1e-3) / 1e12;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (sink) (void)hipFree(sink);
  if (b) (void)hipFree(b);
  if (a) (void)hipFree(a);
  return e;
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
