// Shared definitions of the gfx950 kernels and the engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(4))) float f32x4_t;
typedef __attribute__((ext_vector_type(16))) float f32x16_t;
typedef __attribute__((ext_vector_type(4))) double f64x4_t;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4_t;

namespace lri {

// Device-side status bits latched into ctx->flags (read at synchronising calls).
enum : int { FLAG_BREAKDOWN = 1, FLAG_NONFINITE = 2 };

constexpr int kMaxRank = 64;       // state rank capacity
constexpr int kMaxAug = 2 * kMaxRank;  // rank-adaptive augmented width

// ---------------------------------------------------------------- snapshot decode
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

__device__ __forceinline__ double snap_f64(const double* p, int64_t k) { return p[k]; }
__device__ __forceinline__ double snap_f64(const float* p, int64_t k) { return double(p[k]); }
__device__ __forceinline__ double snap_f64(const uint16_t* p, int64_t k) { return double(bf16_to_f32(p[k])); }

// Round-to-nearest-even float -> bfloat16 bits (finite inputs).
__device__ __forceinline__ uint32_t f32_to_bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}

// Split x into hi + lo, both bfloat16 (returned as bit patterns); |x - hi - lo| <= 2^-18 |x|.
__device__ __forceinline__ void split_bf16(float x, uint32_t& hi, uint32_t& lo) {
  hi = f32_to_bf16_rne(x);
  float r = x - __uint_as_float(hi << 16);
  lo = f32_to_bf16_rne(r);
}

// ---------------------------------------------------------------- small products on the matrix cores
// C(m, n) = sum_kk A(b, m, kk) B(b, kk, n) for b < nbatch, in 16 x 16 blocks
// distributed over the workgroup's waves (v_mfma_f64_16x16x4_f64: lane l holds
// A(m0 + l % 16, k0 + l / 16), B(k0 + l / 16, n0 + l % 16) and the results
// (m0 + 4 u + l / 16, n0 + l % 16), u < 4); store(b, m, n, value) is called once
// per element.  The operand functors are only called in range.  Every wave of
// the workgroup must call it (the loop bounds are wave-uniform).
template <class FA, class FB, class FS>
__device__ __forceinline__ void wave_gemm(int M, int N, int K, int nbatch, FA A, FB B, FS store) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int tm = (M + 15) >> 4, tn = (N + 15) >> 4;
  const int per = tm * tn;
  for (int tile = wave; tile < per * nbatch; tile += nw) {
    const int bt = tile / per, rem = tile - bt * per;
    const int m0 = (rem % tm) << 4, n0 = (rem / tm) << 4;
    const int mm = m0 + lr, nn = n0 + lr;
    const int mc = mm < M ? mm : M - 1, nc = nn < N ? nn : N - 1;  // in-range rows for the loads
    // two accumulators (even and odd k-steps): a dependent FP64 matrix-core
    // product waits ~190 cycles for the previous one, so this halves the chain
    f64x4_t acc = {0.0, 0.0, 0.0, 0.0}, acc2 = {0.0, 0.0, 0.0, 0.0};
    int k0 = 0;
    // four k-steps per round: the eight operand loads are unconditional (at
    // clamped indices, zeroed by selects) and issue together
    for (; k0 + 16 <= K; k0 += 16) {
      double a[4], b[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        a[s] = A(bt, mc, k0 + 4 * s + lk);
        b[s] = B(bt, k0 + 4 * s + lk, nc);
      }
#pragma unroll
      for (int s = 0; s < 4; s += 2) {
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(mm < M ? a[s] : 0.0, nn < N ? b[s] : 0.0, acc, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(mm < M ? a[s + 1] : 0.0, nn < N ? b[s + 1] : 0.0, acc2, 0, 0,
                                                    0);
      }
    }
    for (; k0 < K; k0 += 4) {
      const int kk = k0 + lk;
      const double a = (mm < M && kk < K) ? A(bt, mm, kk) : 0.0;
      const double b = (kk < K && nn < N) ? B(bt, kk, nn) : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    acc += acc2;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int mo = m0 + 4 * u + lk;
      if (mo < M && nn < N) store(bt, mo, nn, acc[u]);
    }
  }
}

// ---------------------------------------------------------------- packed operand layouts
// FP64 products take the B operand row-major (k x rb, rb a multiple of 16);
// MIXED products take column-major bf16 hi/lo (or f32) panels, c-major with a
// leading dimension padded to a multiple of 64.
inline int64_t pad_to(int64_t x, int64_t q) { return (x + q - 1) / q * q; }

}  // namespace lri
