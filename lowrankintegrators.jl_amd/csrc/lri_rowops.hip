// FP64 kernels on the thin n x r (m x r) factors and the r x r cores.
//
// k_rows is row-parallel over the tall dimension (tiles of 16..256 rows,
// grid-stride over tiles); its products run on FP64 matrix cores and it may
// end in a global reduction C = out' * W (or the Gram out' * out), written as
// one partial per workgroup.  k_reduce sums the partials in a fixed order (so
// results are bitwise reproducible) and k_epilogue runs a small epilogue on C
// in one workgroup: a Cholesky factorisation for CholeskyQR2, the S-step
// update, the Galerkin matrices of the BUG steps.
//
// These replace the LAPACK/BLAS calls of the reference on the factors:
// qr!/Matrix(Q) (projector_splitting.jl:137-138,149-150,174-175,186-187;
// unconventional.jl:141-152; rank_adaptive_unconventional.jl:204-217), the
// U*S / V*S' products (projector_splitting.jl:133,145,170,182), and the r x r
// S-step algebra (projector_splitting.jl:141-142; unconventional.jl:154-156;
// rank_adaptive_unconventional.jl:219-231).
#include <type_traits>

#include "lri_common.h"
#include "lri_kernels.h"

namespace lri {

constexpr int kThreads = 256;
constexpr double kEps = 2.220446049250313e-16;

// ----------------------------------------------------------------------------
// small dense routines on LDS-resident matrices (column-major, ld = k)
// ----------------------------------------------------------------------------

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// 1/sqrt(d) for normal d > 0: the hardware estimate refined by two Newton
// steps y <- y + y (1 - d y^2) / 2.
__device__ __forceinline__ double rsqrt_nr(double d) {
  double y = __builtin_amdgcn_rsq(d);
  double r = __builtin_fma(-d * y, y, 1.0);
  y = __builtin_fma(0.5 * y, r, y);
  r = __builtin_fma(-d * y, y, 1.0);
  return __builtin_fma(0.5 * y, r, y);
}

// 1/d for normal d: the hardware estimate refined by two Newton steps
// y <- y + y (1 - d y).
__device__ __forceinline__ double rcp_nr(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double r = __builtin_fma(-d, y, 1.0);
  y = __builtin_fma(y, r, y);
  r = __builtin_fma(-d, y, 1.0);
  return __builtin_fma(y, r, y);
}

// Exponent e (even) with max_diag * 2^-e in [0.5, 2): the Gram matrix is
// scaled by 2^-e (exactly), so R scales by 2^(-e/2) and every pivot stays a
// normal number; zero, NaN and Inf give 0.
__device__ __forceinline__ int chol_scale_exp(double md) {
  int ex = 0;
  if (md > 0.0 && isfinite(md)) frexp(md, &ex);
  return ex & ~1;
}

// Upper Cholesky G = R'R (k <= K <= 64) and R^{-1}, right-looking on the
// augmented system [W | B], W the scaled G and B = L^{-1} (L = R', B starting
// as I): one elimination sweep yields R and R^{-T} together, with no back
// substitution.  G is scaled by 2^-e (exactly; max_i G(i,i) 2^-e in [0.5, 2))
// so that every pivot stays a normal number, and R and R^{-1} are unscaled at
// the end.  A pivot <= tau = 10 eps max_i G(i,i) (scaled) marks column p as
// numerically dependent: dropped[p] = 1, R(p,p) = 1, row p of R is zero and
// row and column p of R^{-1} are zero (so Q gets a zero column there).  R and
// Rinv are written to LDS (column-major, ld k).  Returns the number of
// dropped columns; sets *nonfinite on NaN/Inf.  Called by all threads.
//
// The rows of [W | B] are spread over the first four waves in blocks of four
// (wave w: row blocks w, w + 4, ...; at group g of 16 pivots, slot s of a
// wave's register window holds row 16 (g + s / 4) + 4 w + s % 4, and the
// window shifts by one block after each group) and the columns over the lanes
// (lane j: column j of W and of B).  The wave owning a block factors its four
// pivots in registers, the pivot rows exchanged by readlane, writes the four
// rows R(p, :) and L^{-1}(p, :) into one of two LDS buffers, and after one
// barrier every other wave applies them to its rows (K/4 slots, 2 FMAs each).
// Finished rows read zero multipliers, so no row needs a mask, and the group
// loop is not unrolled.  At k = 32 this takes ~26K cycles, against ~35K with
// one barrier per pivot and ~40K for one wave unrolled over the pivots (LDS
// broadcasts).  G is not modified.
template <int K>
__device__ int chol_inv_waves(const double* __restrict__ G, double* __restrict__ R, double* __restrict__ Rinv, int k,
                              unsigned char* dropped, int* nonfinite) {
  static_assert(K == 16 || K == 32 || K == 64, "K: 16, 32 or 64");
  constexpr int NS = K / 4;
  __shared__ double s_rowR[2][4][128];  // R(p, j) by column j; [K, 128) stays zero (rows past the window)
  __shared__ double s_rowB[2][4][64];   // L^{-1}(p, j)
  __shared__ int s_nfw[4];
  const int t = threadIdx.x, wv = t >> 6, j = t & 63;
  for (int e = t; e < 2 * 4 * 128; e += blockDim.x) (&s_rowR[0][0][0])[e] = 0.0;
  // the scale, found by every wave
  double md = j < k ? G[j + j * k] : 0.0;
  int nf0 = (j < k) && !isfinite(md);
  for (int off = 32; off > 0; off >>= 1) {
    md = fmax(md, __shfl_xor(md, off));
    nf0 |= __shfl_xor(nf0, off);
  }
  const int e2 = chol_scale_exp(md), eh = e2 / 2;
  const double tau = 10.0 * kEps * ldexp(md, -e2);
  double a[NS], b[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int tt = 16 * (s / 4) + 4 * wv + (s & 3);
    a[s] = (wv < 4 && tt < k && j < k && tt <= j) ? ldexp(G[tt + j * k], -e2) : 0.0;
    b[s] = (wv < 4 && tt == j && j < k) ? 1.0 : 0.0;
  }
  int nfl = nf0;
  __syncthreads();
  int buf = 0;
  for (int g = 0; 16 * g < k; ++g) {
#pragma unroll
    for (int wp = 0; wp < 4; ++wp) {
      const int p0 = 16 * g + 4 * wp;
      if (p0 < k) {
        if (wv == wp) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int p = p0 + i;
            if (p < k) {
              const double d = readlane_f64(a[i], p);
              const bool fin = isfinite(d);
              const bool keep = (d > tau) && fin;
              const double inv0 = rsqrt_nr(keep ? d : 1.0);
              const double inv = keep ? inv0 : 1.0;
              const double mR = (keep && j > p && j < k) ? a[i] * inv : 0.0;
              const double mB = (keep && j < k) ? b[i] * inv : 0.0;
#pragma unroll
              for (int i2 = i + 1; i2 < 4; ++i2) {
                const double rt = readlane_f64(mR, (p0 + i2) & 63);
                a[i2] = __builtin_fma(-rt, mR, a[i2]);
                b[i2] = __builtin_fma(-rt, mB, b[i2]);
              }
              s_rowR[buf][i][j] = mR;
              s_rowB[buf][i][j] = mB;
              if (j < k) {
                R[p + j * k] = j == p ? (keep ? d * inv : 1.0) : mR;
                Rinv[j + p * k] = mB;
              }
              if (j == p) dropped[p] = keep ? 0 : 1;
              nfl |= !fin;
            } else {
              s_rowR[buf][i][j] = 0.0;
              s_rowB[buf][i][j] = 0.0;
            }
          }
        }
        __syncthreads();
        if (wv < 4) {
          // the owner's own block is done; its later blocks start at slot 4
          const int s0 = wv == wp ? 4 : 0;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const double mRj = s_rowR[buf][i][j], mBj = s_rowB[buf][i][j];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
              if (s >= s0) {
                const double rt = s_rowR[buf][i][16 * (g + s / 4) + 4 * wv + (s & 3)];
                a[s] = __builtin_fma(-rt, mRj, a[s]);
                b[s] = __builtin_fma(-rt, mBj, b[s]);
              }
            }
          }
        }
        buf ^= 1;
      }
    }
#pragma unroll
    for (int s = 0; s + 4 < NS; ++s) {
      a[s] = a[s + 4];
      b[s] = b[s + 4];
    }
#pragma unroll
    for (int s = (NS > 4 ? NS - 4 : 0); s < NS; ++s) {
      a[s] = 0.0;
      b[s] = 0.0;
    }
  }
  if (j == 0 && wv < 4) s_nfw[wv] = nfl;
  __syncthreads();
  // unscale; the rows of R^{-1} of dropped columns are zero (lanes: rows, waves: columns)
  const int nw = blockDim.x >> 6;
  if (j < k)
    for (int c = wv; c < k; c += nw) {
      R[j + c * k] = ldexp(R[j + c * k], eh);
      Rinv[j + c * k] = dropped[j] ? 0.0 : ldexp(Rinv[j + c * k], -eh);
    }
  const int nd = __popcll(__ballot(j < k && dropped[j]));
  if (t == 0 && (s_nfw[0] | s_nfw[1] | s_nfw[2] | s_nfw[3])) *nonfinite = 1;
  __syncthreads();
  return nd;
}

// The same factorisation (same scaling, drop rule and outputs) for k <= K <= 32
// on one wave, blocked by four pivots, with the trailing updates on the FP64
// matrix cores.  The whole scaled G and the running B = L^{-1} stay in
// registers in the v_mfma_f64_16x16x4_f64 result layout: 16 x 16 tiles (T1, T2),
// lane l and item u holding row 16 T1 + 4 u + l / 16, column 16 T2 + l % 16.
// Since G is kept symmetric, the panel of pivot block b (columns 4b .. 4b+3)
// in the operand layout of that instruction (lane l: row 16 t + l % 16, column
// 4b + l / 16) is item b % 4 of tile (b / 4, t) at the same lane: no data moves
// to form it.  Per block:
//  * the 4 x 4 diagonal block D is read by readlane (lane 16 q + 4 (b % 4) + a
//    holds D(a, q)) and factored in uniform registers with the drop rule
//    (keep_q = pivot > tau; inv_q = 1/sqrt(pivot)), giving the multipliers
//    Lb(a, q) and the upper triangular M with L(i, 4b + q) = sum_p A(i, 4b + p) M(p, q):
//    M(:, q) = keep_q inv_q (e_q - sum_{p<q} M(:, p) Lb(q, p));
//  * the panel L = A M takes four cross-lane reads per tile (ds_bpermute);
//  * B's block row becomes M' times itself (rows 4b + a live in item b % 4 of
//    tiles (b / 4, .) at lanes with l / 16 = a);
//  * G -= P P' and B -= P B(block row) on the matrix cores, P the panel with
//    its rows <= 4b + 3 zeroed (finished rows and columns stay untouched).
// The four pivots of a block form the only sequential chain: ~1,000 cycles per
// block against ~770 per pivot for the row-wise forms.  All threads call it;
// the first wave computes, the others wait at the final barrier.
template <int K>
__device__ int chol_inv_mfma(const double* __restrict__ G, double* __restrict__ R, double* __restrict__ Rinv, int k,
                             unsigned char* dropped, int* nonfinite) {
  static_assert(K == 16 || K == 32, "K: 16 or 32");
  constexpr int NT = K / 16, NB = K / 4;
  const int t = threadIdx.x, j = t & 63;
  if (t < 64) {
    const int lr = j & 15, lk = j >> 4;
    double md = j < k ? G[j + j * k] : 0.0;
    int nf0 = (j < k) && !isfinite(md);
    for (int off = 32; off > 0; off >>= 1) {
      md = fmax(md, __shfl_xor(md, off));
      nf0 |= __shfl_xor(nf0, off);
    }
    const int e2 = chol_scale_exp(md), eh = e2 / 2;
    const double tau = 10.0 * kEps * ldexp(md, -e2);
    f64x4_t g[NT][NT], x[NT][NT];
#pragma unroll
    for (int T1 = 0; T1 < NT; ++T1)
#pragma unroll
      for (int T2 = 0; T2 < NT; ++T2)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int row = 16 * T1 + 4 * u + lk, col = 16 * T2 + lr;
          const bool in = row < k && col < k;
          const int lo = row < col ? row : col, hi = row < col ? col : row;  // the upper triangle
          const double v = G[in ? lo + hi * k : 0];
          g[T1][T2][u] = in ? ldexp(v, -e2) : 0.0;
          x[T1][T2][u] = row == col ? 1.0 : 0.0;
        }
    // one-hot weights of this lane's block column q = l / 16 (selects of uniform
    // values by lane, as multiply-adds: lane-dependent selects compile to branches)
    double w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = lk == q ? 1.0 : 0.0;
    // cross-lane read addresses: lane 16 p + l % 16
    int src[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) src[p] = (16 * p + lr) << 2;
    auto bperm = [](int addr, double v) {
      const uint64_t u = __builtin_bit_cast(uint64_t, v);
      const uint32_t lo = __builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)u);
      const uint32_t hi = __builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)(u >> 32));
      return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
    };
    // pivots past k are zero and count as dropped: no mask is needed in the sweep
    uint32_t dm = 0;
    int nfv = 0;
    double rR[NB][NT];
#pragma unroll
    for (int kb = 0; kb < NB; ++kb) {
      const int T = kb >> 2, u0 = kb & 3;
      double P[NT];
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) P[tt] = g[T][tt][u0];
      double W[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int q = 0; q <= a; ++q) W[a][q] = readlane_f64(P[T], 16 * q + 4 * u0 + a);
      // the block's four pivots; a dropped pivot gets inv = 0, which zeroes its
      // multipliers and its column of M
      double Lb[4][4], inv[4], dq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double d = W[q][q];
        const bool keep = d > tau && d <= 1.7976931348623157e308;
        nfv |= !(fabs(d) <= 1.7976931348623157e308);
        dm |= keep ? 0u : 1u << (4 * kb + q);
        inv[q] = keep ? rsqrt_nr(d) : 0.0;
        dq[q] = keep ? 0.0 : 1.0;
#pragma unroll
        for (int a = q + 1; a < 4; ++a) Lb[a][q] = W[a][q] * inv[q];
#pragma unroll
        for (int a = q + 1; a < 4; ++a)
#pragma unroll
          for (int b = q + 1; b <= a; ++b) W[a][b] = __builtin_fma(-Lb[a][q], Lb[b][q], W[a][b]);
      }
      double M[4][4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double c[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) c[i] = i == q ? 1.0 : 0.0;
#pragma unroll
        for (int p = 0; p < q; ++p)
#pragma unroll
          for (int i = 0; i <= p; ++i) c[i] = __builtin_fma(-M[i][p], Lb[q][p], c[i]);
#pragma unroll
        for (int i = 0; i < 4; ++i) M[i][q] = i <= q ? c[i] * inv[q] : 0.0;
      }
      double mq[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        double v = M[p][3] * w[3];
#pragma unroll
        for (int q = 2; q >= p; --q) v = __builtin_fma(M[p][q], w[q], v);
        mq[p] = v;
      }
      const double drq = __builtin_fma(dq[0], w[0], __builtin_fma(dq[1], w[1], __builtin_fma(dq[2], w[2], dq[3] * w[3])));
      double Pm[NT];
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        double acc = bperm(src[0], P[tt]) * mq[0];
#pragma unroll
        for (int p = 1; p < 4; ++p) acc = __builtin_fma(bperm(src[p], P[tt]), mq[p], acc);
        // R(p, row) = L(row, p), p = 4 kb + l / 16: the panel at and below the
        // block's diagonal (1 on the diagonal of a dropped pivot), zero above
        const int row = 16 * tt + lr, a = row - 4 * kb;
        const double rv = a > lk ? acc : 0.0;
        rR[kb][tt] = a == lk ? acc + drq : rv;
        Pm[tt] = row >= 4 * kb + 4 ? acc : 0.0;
      }
      double xn[NT];
#pragma unroll
      for (int t2 = 0; t2 < NT; ++t2) {
        double acc = bperm(src[0], x[T][t2][u0]) * mq[0];
#pragma unroll
        for (int b = 1; b < 4; ++b) acc = __builtin_fma(bperm(src[b], x[T][t2][u0]), mq[b], acc);
        xn[t2] = acc;
      }
#pragma unroll
      for (int t2 = 0; t2 < NT; ++t2) x[T][t2][u0] = xn[t2];
#pragma unroll
      for (int t1 = T; t1 < NT; ++t1) {
#pragma unroll
        for (int t2 = T; t2 < NT; ++t2)
          g[t1][t2] = __builtin_amdgcn_mfma_f64_16x16x4f64(-Pm[t1], Pm[t2], g[t1][t2], 0, 0, 0);
#pragma unroll
        for (int t2 = 0; t2 <= T; ++t2)
          x[t1][t2] = __builtin_amdgcn_mfma_f64_16x16x4f64(-Pm[t1], xn[t2], x[t1][t2], 0, 0, 0);
      }
    }
    // R unscaled; R^{-1} = B' unscaled, the rows of dropped columns zero
#pragma unroll
    for (int kb = 0; kb < NB; ++kb)
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        const int p = 4 * kb + lk, row = 16 * tt + lr;
        if (p < k && row < k) R[p + row * k] = ldexp(rR[kb][tt], eh);
      }
#pragma unroll
    for (int T1 = 0; T1 < NT; ++T1)
#pragma unroll
      for (int T2 = 0; T2 < NT; ++T2)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int c = 16 * T1 + 4 * u + lk, jj = 16 * T2 + lr;
          if (c < k && jj < k) Rinv[jj + c * k] = ((dm >> jj) & 1) ? 0.0 : ldexp(x[T1][T2][u], -eh);
        }
    if (j < k) dropped[j] = (dm >> j) & 1;
    // a non-finite pivot of a row past k cannot occur (those rows are zero)
    if (t == 0 && (nfv | nf0)) *nonfinite = 1;
  }
  __syncthreads();
  return __popcll(__ballot(j < k && dropped[j]));
}

// CK: the Cholesky variant compiled into an epilogue kernel (16, 32, 64: the
// bound on k): the one-wave matrix-core form up to 32, the four-wave form above.
template <int CK>
__device__ int chol_inv(const double* G, double* R, double* Rinv, int k, unsigned char* dropped, int* nonfinite) {
  if constexpr (CK <= 32) return chol_inv_mfma<CK>(G, R, Rinv, k, dropped, nonfinite);
  else return chol_inv_waves<CK>(G, R, Rinv, k, dropped, nonfinite);
}

// C = op(A) * op(B), k x k, column-major, on the FP64 matrix cores (16 x 16
// blocks over the workgroup's waves).  C must not alias A or B; every thread
// of the workgroup calls it.
__device__ void small_gemm(const double* __restrict__ A, int tA, const double* __restrict__ B, int tB,
                           double* __restrict__ C, int k) {
  // (the transposes select the index, not between two loads: a select of
  // loads compiles to branches)
  wave_gemm(
      k, k, k, 1, [&](int, int i, int l) { return A[tA ? l + i * k : i + l * k]; },
      [&](int, int l, int j) { return B[tB ? j + l * k : l + j * k]; },
      [&](int, int i, int j, double v) { C[i + j * k] = v; });
}

// LDS doubles the epilogue needs beyond C (a x b).
__host__ __device__ inline int epi_scratch_doubles(int kind, int a, int b) {
  int k = a > b ? a : b;
  switch (kind) {
    case EPI_CHOL1:
    case EPI_CHOL2:
    case EPI_CHOL3:
      return 3 * a * a;  // R, R^{-1} and Rtot (C may carry a cross term beyond its first a columns)
    case EPI_BUG_N:
      return 2 * k * k;
    case EPI_RA_N:
      return 2 * k * k;
    default:
      return 0;
  }
}

// ----------------------------------------------------------------------------
// epilogues run by the last workgroup on the reduced C (LDS, a x b)
// ----------------------------------------------------------------------------

// writer: this workgroup stores the results to the state (the epilogue of a
// reduction runs in one workgroup; folded into the prologue of the next row
// kernel, every workgroup computes the factors for itself and workgroup 0
// stores them).  The Cholesky kinds use C's first a columns as the Gram matrix.
template <int CK>
__device__ void run_epilogue(const EpiArgs& E, double* C, int a, int b, double* scr, bool writer = true) {
  const int t = threadIdx.x;
  SmallState* st = E.small;
  switch (E.kind) {
    case EPI_CHOL1:
    case EPI_CHOL2:
    case EPI_CHOL3: {
      const int k = a;
      __shared__ unsigned char dropped[kMaxAug];
      __shared__ int s_big;
      __shared__ double s_shift;
      if (E.kind == EPI_CHOL1 && E.shift_rows > 0) {
        // shifted CholeskyQR3, first factor: G + s I (Fukaya, Kannan, Nakatsukasa, Yamamoto,
        // Yanagisawa 2020), with ||G||_2 bounded by trace(G)
        if (t < 64) {
          double tr = 0.0;
          for (int i = t; i < k; i += 64) tr += C[i + i * k];
          for (int off = 32; off > 0; off >>= 1) tr += __shfl_xor(tr, off);
          if (t == 0) s_shift = 11.0 * ((double)E.shift_rows * k + (double)k * (k + 1)) * (0.5 * kEps) * tr;
        }
        __syncthreads();
        for (int i = t; i < k; i += blockDim.x) C[i + i * k] += s_shift;
        __syncthreads();
      }
      double* R = scr;
      double* Rinv = scr + k * k;
      int nf = 0, nd = 0;
      // CHOL2/3: the factor on the right of Rtot (R1, or the running Rtot) and
      // the first pass's dropped count were written by an earlier launch on
      // this stream; their loads are issued now, ahead of the factorisation
      // (for CK <= 32; the CK = 64 variant has no registers to spare)
      constexpr int NQ = CK <= 32 ? (CK * CK) / kThreads : 0;
      const double* right = E.kind == EPI_CHOL2 ? st->R1 : st->Rtot;
      double rv[NQ > 0 ? NQ : 1];
      if (E.kind != EPI_CHOL1) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int e = t + kThreads * q;
          rv[q] = e < k * k ? right[e] : 0.0;
        }
      }
      const int nd_first = E.kind != EPI_CHOL1 ? st->ndropped : 0;
      // CholeskyQR2's second Gram matrix is I + E with |E| ~ eps cond^2: when
      // max|E| <= 1e-6 (and the first pass dropped nothing), R = I + F with
      // F = F1 + F2, F1 = upper(E) (diagonal halved), F2 = upper(-F1'F1)
      // (diagonal halved), and R^{-1} = I - F + F^2, all exact to O(|E|^3).
      bool pert = false;
      if (E.kind != EPI_CHOL1 && nd_first == 0) {
        if (t == 0) s_big = 0;
        __syncthreads();
        double mx = 0.0;
        for (int e = t; e < k * k; e += blockDim.x) {
          const int i = e % k, jj = e / k;
          const double v = C[e] - (i == jj ? 1.0 : 0.0);
          mx = fmax(mx, isfinite(v) ? fabs(v) : 1e300);
        }
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        if ((t & 63) == 0 && mx > 1e-6) atomicOr(&s_big, 1);  // any wave over the bound -> full Cholesky
        __syncthreads();
        pert = (s_big == 0);
      }
      if (pert) {
        double* F1 = Rinv;  // scratch: F1, then F1'F1 into R
        for (int e = t; e < k * k; e += blockDim.x) {
          const int i = e % k, jj = e / k;
          const double v = C[e] - (i == jj ? 1.0 : 0.0);
          F1[e] = i < jj ? v : (i == jj ? 0.5 * v : 0.0);
        }
        __syncthreads();
        small_gemm(F1, 1, F1, 0, R, k);  // R <- F1' F1
        __syncthreads();
        for (int e = t; e < k * k; e += blockDim.x) {  // C <- F = F1 + upper(-F1'F1) (diag halved)
          const int i = e % k, jj = e / k;
          const double f2 = i < jj ? -R[e] : (i == jj ? -0.5 * R[e] : 0.0);
          C[e] = F1[e] + f2;
        }
        __syncthreads();
        small_gemm(C, 0, C, 0, Rinv, k);  // Rinv <- F^2
        __syncthreads();
        for (int e = t; e < k * k; e += blockDim.x) {
          const int i = e % k, jj = e / k;
          const double id = i == jj ? 1.0 : 0.0;
          R[e] = id + C[e];
          Rinv[e] = id - C[e] + Rinv[e];
        }
        if (t < k) dropped[t] = 0;
        __syncthreads();
      } else {
        nd = chol_inv<CK>(C, R, Rinv, k, dropped, &nf);
      }
      if (E.kind == EPI_CHOL1) {
        if (writer)
          for (int e = t; e < k * k; e += blockDim.x) {
            st->R1[e] = R[e];
            st->R1inv[e] = Rinv[e];
          }
      } else {
        // Rtot = R2 * R1 (CHOL2) or R3 * Rtot (CHOL3) into the third scratch
        // matrix; the factor on the right was written by an earlier launch on
        // this stream; C is free now
        double* Rt = scr + 2 * k * k;
        if (writer)
          for (int e = t; e < k * k; e += blockDim.x) st->R2inv[e] = Rinv[e];
        if (NQ > 0) {
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const int e = t + kThreads * q;
            if (e < k * k) C[e] = rv[q];
          }
        } else {
          for (int e = t; e < k * k; e += blockDim.x) C[e] = right[e];
        }
        __syncthreads();
        small_gemm(R, 0, C, 0, Rt, k);
        __syncthreads();
        if (writer)
          for (int e = t; e < k * k; e += blockDim.x) {
            const int i = e % k, jj = e / k;
            st->Rtot[e] = Rt[e];
            if (E.S) E.S[i + jj * E.ld_s] = E.transpose ? Rt[jj + i * k] : Rt[e];
          }
        if (E.stilde) {
          // C = [Q1'Q1 | Q1'W]: U1 = Q1 R2^{-1}, so U1'W = R2^{-T} (Q1'W) and
          // S~ = Rtot - U1'W; the first k columns of C are free again
          small_gemm(Rinv, 1, C + k * k, 0, C, k);
          __syncthreads();
          if (writer)
            for (int e = t; e < k * k; e += blockDim.x) {
              const int i = e % k, jj = e / k;
              const double v = Rt[e] - C[e];
              if (E.stilde == 2) st->St[jj + i * k] = v;
              else st->St[e] = v;
            }
        }
      }
      if (t == 0 && writer) {
        if (nf) atomicOr(E.flags, FLAG_NONFINITE);
        if (nd && E.breakdown_is_error) atomicOr(E.flags, FLAG_BREAKDOWN);
        if (E.kind == EPI_CHOL1) st->ndropped = nd;
      }
      break;
    }
    case EPI_STILDE: {
      const int k = a;
      for (int e = t; e < k * k; e += blockDim.x) {
        int i = e % k, j = e / k;
        double v = st->Rtot[e] - C[e];
        if (E.transpose) st->St[j + i * k] = v;
        else st->St[e] = v;
      }
      break;
    }
    case EPI_SFINAL: {
      const int k = a;
      for (int e = t; e < k * k; e += blockDim.x) {
        int i = e % k, j = e / k;
        E.S[i + j * E.ld_s] = E.transpose ? st->Rtot[j + i * k] : st->Rtot[e];
      }
      break;
    }
    case EPI_BUG_M: {
      for (int e = t; e < a * b; e += blockDim.x) st->M[e] = C[e];
      break;
    }
    case EPI_BUG_N: {
      // N = C; Sinit = M (S0 N') with S0, then M, staged in the first scratch matrix
      const int k = a;
      double* W = scr;
      double* T = scr + k * k;
      for (int e = t; e < k * k; e += blockDim.x) {
        st->N[e] = C[e];
        W[e] = E.S[(e % k) + (e / k) * E.ld_s];
      }
      __syncthreads();
      small_gemm(W, 0, C, 1, T, k);  // T = S0 N'
      __syncthreads();
      for (int e = t; e < k * k; e += blockDim.x) W[e] = st->M[e];
      __syncthreads();
      small_gemm(W, 0, T, 0, st->Sinit, k);
      break;
    }
    case EPI_BUG_S: {
      const int k = a;
      for (int e = t; e < k * k; e += blockDim.x) {
        int i = e % k, j = e / k;
        E.S[i + j * E.ld_s] = st->Sinit[e] + C[j + i * k];
      }
      break;
    }
    case EPI_RA_M:
    case EPI_RA_N: {
      const int r = a, k2 = 2 * r;
      double* G = (E.kind == EPI_RA_M) ? st->M : st->N;
      for (int e = t; e < k2 * r; e += blockDim.x) {
        int i = e % k2, j = e / k2;
        G[e] = (i < r) ? (i == j ? 1.0 : 0.0) : C[(i - r) + j * r];
      }
      if (E.kind == EPI_RA_M) break;
      __syncthreads();
      double* T = scr;  // M * S0, 2r x r
      for (int e = t; e < k2 * r; e += blockDim.x) {
        int i = e % k2, j = e / k2;
        double s = 0.0;
        for (int l = 0; l < r; ++l) s += st->M[i + l * k2] * E.S[l + j * E.ld_s];
        T[e] = s;
      }
      __syncthreads();
      for (int e = t; e < k2 * k2; e += blockDim.x) {
        int i = e % k2, j = e / k2;
        double s = 0.0;
        for (int l = 0; l < r; ++l) s += T[i + l * k2] * st->N[j + l * k2];
        st->Sinit[e] = s;
      }
      break;
    }
    case EPI_COPYLTU: {
      for (int e = t; e < a * a; e += blockDim.x) {
        const int i = e % a, j = e / a;
        const int p = i >= j ? i : j, q = i >= j ? j : i;  // the lower-triangle entry (p >= q)
        E.dst[e] = (E.S ? E.S[p + q * E.ld_s] : 0.0) - C[p + q * a];
      }
      break;
    }
    case EPI_STORE: {
      for (int e = t; e < a * b; e += blockDim.x) st->Ctmp[e] = C[e];
      break;
    }
    case EPI_COPY: {
      const int ld = E.ld_dst > 0 ? E.ld_dst : (E.transpose ? b : a);
      for (int e = t; e < a * b; e += blockDim.x) {
        const int i = e % a, j = e / a;
        if (E.transpose)
          E.dst[j + i * ld] = C[e];
        else
          E.dst[i + j * ld] = C[e];
      }
      break;
    }
    case EPI_RA_S: {
      const int k2 = a;
      for (int e = t; e < k2 * k2; e += blockDim.x) {
        int i = e % k2, j = e / k2;
        st->Shat[e] = st->Sinit[e] + C[j + i * k2];
      }
      break;
    }
    default:
      break;
  }
}

// Partial sums of the per-workgroup reductions: workgroup b of k_reduce sums
// entries [64b, 64b + 64) of the nb partials (each a*b doubles) into C; thread
// (entry e, quarter g) adds partials g, g+4, ... with 16 loads in flight, and the
// four quarters are combined in a fixed order (deterministic).
__device__ __forceinline__ void reduce_entries(const double* __restrict__ gpart, int nb, int ab,
                                               double* __restrict__ C, int bid) {
  __shared__ double q4[4][64];
  const int el = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int e = bid * 64 + el;
  double s = 0.0;
  if (e < ab) {
    int k = g;
    for (; k + 4 * 15 < nb; k += 4 * 16) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = gpart[(int64_t)(k + 4 * u) * ab + e];
#pragma unroll
      for (int u = 0; u < 16; ++u) s += v[u];
    }
    for (; k < nb; k += 4) s += gpart[(int64_t)k * ab + e];
  }
  q4[g][el] = s;
  __syncthreads();
  if (g == 0 && e < ab) C[e] = ((q4[0][el] + q4[1][el]) + q4[2][el]) + q4[3][el];
}

__global__ __launch_bounds__(kThreads) void k_reduce(const double* __restrict__ gpart, int nb, int ab,
                                                     double* __restrict__ C) {
  reduce_entries(gpart, nb, ab, C, blockIdx.x);
}

struct ReduceLaunch {
  const double* gpart;
  int nb, ab;
  double* C;
};
__global__ __launch_bounds__(kThreads) void k_reduce_batch(const unsigned char* __restrict__ tab,
                                                           const int* __restrict__ start, int count) {
  const int b = blockIdx.x;
  const int i = batch_member(start, count, b);
  const ReduceLaunch& R = *reinterpret_cast<const ReduceLaunch*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes);
  reduce_entries(R.gpart, R.nb, R.ab, R.C, b - start[i]);
}

// One workgroup: C (a x b, from k_reduce) into LDS, then the epilogue.  (A
// variant where the last k_reduce workgroup ran the epilogue after a
// device-scope release/acquire measured slower than this second launch.)
template <int CK>
__device__ __forceinline__ void epilogue_body(const EpiArgs& E, const double* __restrict__ Cg, int a, int b) {
  extern __shared__ double lds[];
  double* C = lds;
  double* scr = lds + a * b;
  for (int e = threadIdx.x; e < a * b; e += kThreads) C[e] = Cg[e];
  __syncthreads();
  run_epilogue<CK>(E, C, a, b, scr);
}
template <int CK>
__global__ __launch_bounds__(kThreads) void k_epilogue(EpiArgs E, const double* __restrict__ Cg, int a, int b) {
  epilogue_body<CK>(E, Cg, a, b);
}

struct EpiLaunch {
  EpiArgs E;
  const double* Cg;
  int a, b;
};
template <int CK>
__global__ __launch_bounds__(kThreads) void k_epilogue_batch(const unsigned char* __restrict__ tab,
                                                             const int* __restrict__ start, int count) {
  const int i = batch_member(start, count, blockIdx.x);
  const EpiLaunch& L = *reinterpret_cast<const EpiLaunch*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes);
  epilogue_body<CK>(L.E, L.Cg, L.a, L.b);
}


// ----------------------------------------------------------------------------
// the row kernel
//   out[i, c] = (parts ? sum_s part[s][c][i] : 0) + base_sign * op(base)[i, c]
//             + xy_sign * sum_d X[i, d] * op(Y)(d, c)
// and optionally the reduction C = out' * W (W == nullptr: out' * out;
// gram_w: out' * [out | W]).
// A tile has rpb rows.  For loads and stores thread t handles row t % rpb and
// the columns 8*(t / rpb) + 0..7, so every wave reads and writes 64
// consecutive rows of one column.  The tile then goes through LDS (Xt: the X
// rows, Ot: the outputs, Wt: the W rows), where both products run on FP64
// matrix cores (16x16x4): X*Y adds into Ot in place, and the reduction
// accumulates 16x16 blocks of C in registers across tiles (block b belongs to
// wave b % 4), written once per workgroup to gpart[blockIdx.x].
// With a prologue (CK > 0), every workgroup first runs the Cholesky epilogue
// of the previous kernel's reduction on its own LDS copy of C, and Y is that
// epilogue's R^{-1}.
// ----------------------------------------------------------------------------

struct RowsPlan {
  int tpr, rpb;                // threads per row, rows per tile
  int ri4, ro16, rw16;         // padded widths
  int rg16;                    // the out' * out block columns of a gram_w reduction (else 0)
  int ldy, ldxt, ldot, ldwt;   // LDS leading dimensions (doubles; odd against bank conflicts)
  int nxy, nred;               // 16x16 blocks of X*Y per tile, of the reduction
  int ys_off, xt_off, ot_off, wt_off;
  size_t lds_doubles;
};

// threads per row: enough for r_out in groups of 8, and more (fewer rows per
// tile) while the LDS tile would exceed 96 KB (e.g. a narrow output reduced
// against a wide W)
__host__ __device__ inline RowsPlan rows_plan(int ro, int ri, int rw, bool hasX, bool hasW, bool reduce,
                                              bool gramw = false) {
  RowsPlan p;
  p.ri4 = (ri + 3) & ~3;
  p.ro16 = (ro + 15) & ~15;
  p.rw16 = hasW ? ((rw + 15) & ~15) : p.ro16;
  p.rg16 = (hasW && gramw) ? p.ro16 : 0;
  p.ldy = p.ro16 + 1;
  p.ldxt = p.ri4 + 1;
  p.ldot = p.ro16 + 1;
  p.ldwt = p.rw16 + 1;
  int tpr = 1;
  while (8 * tpr < ro) tpr *= 2;
  for (;;) {
    p.tpr = tpr;
    p.rpb = kThreads / tpr;
    int off = 0;
    p.ys_off = off;
    if (hasX) off += p.ri4 * p.ldy;
    p.xt_off = off;
    if (hasX) off += p.rpb * p.ldxt;
    p.ot_off = off;
    if (hasX || reduce) off += p.rpb * p.ldot;
    p.wt_off = off;
    if (hasW && reduce) off += p.rpb * p.ldwt;
    p.lds_doubles = (size_t)off;
    if (p.lds_doubles * sizeof(double) <= 96 * 1024 || tpr >= 16) break;
    tpr *= 2;
  }
  p.nxy = hasX ? (p.rpb / 16) * (p.ro16 / 16) : 0;
  p.nred = reduce ? (p.ro16 / 16) * ((p.rg16 + p.rw16) / 16) : 0;
  return p;
}

// CK > 0: with the prologue of RowArgs::pre (Cholesky variant CK).  bid / nblk:
// this workgroup's index among the nblk of the launch (of its member, batched).
// The arguments are taken by value: bound to a reference into the kernel's
// argument segment, the Cholesky variants spill ~470 bytes per lane to scratch.
// LEAN: fewer loads in flight per thread (2 splits of the partials, 4 X values
// ahead of the Y staging), so that the body fits 256 registers and two
// workgroups share a CU; the arithmetic, and so the result, is the same.
template <typename PT, int NRB, int CK = 0, bool LEAN = false>
__device__ __forceinline__ void rows_body(const RowArgs<PT> A, int bid, int nblk) {
  extern __shared__ double lds[];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int ro = A.r_out, ri = A.r_in;
  const bool hasX = A.X != nullptr;
  const bool hasW = A.W != nullptr;
  const bool reduce = A.reduce != 0;
  const int rw = hasW ? A.rw : ro;
  const RowsPlan p = rows_plan(ro, ri, rw, hasX, hasW, reduce, A.gram_w != 0);
  const int rg = p.rg16 ? ro : 0;  // columns of C before the W part
  const int rpb = p.rpb, tpr = p.tpr;
  const int lrow = t % rpb, cg = t / rpb;
  double* Ys = lds + p.ys_off;
  double* Xt = lds + p.xt_off;
  double* Ot = lds + p.ot_off;
  double* Wt = (hasW && reduce) ? lds + p.wt_off : Ot;
  const int ldwt = (hasW && reduce) ? p.ldwt : p.ldot;
  const bool use_tile = hasX || reduce;
  const int lr = lane & 15, lk = lane >> 4;  // matrix-core lane coordinates

  f64x4_t racc[NRB];
#pragma unroll
  for (int b = 0; b < NRB; ++b) racc[b] = f64x4_t{0.0, 0.0, 0.0, 0.0};

  bool ys_staged = !hasX;
  if constexpr (CK > 0) {
    // the previous reduction's Cholesky epilogue, in the LDS beyond the tile
    const int pa = A.pre_a, pb = A.pre_b;
    double* Cp = lds + p.lds_doubles;
    double* scr = Cp + pa * pb;
    for (int e = t; e < pa * pb; e += kThreads) Cp[e] = A.pre_C[e];
    __syncthreads();
    run_epilogue<CK>(A.pre, Cp, pa, pb, scr, bid == 0);
    __syncthreads();
    // Y = its R^{-1} (pa x pa), staged as below; the tile barrier publishes it
    const double* Ri = scr + pa * pa;
    for (int e = t; e < p.ri4 * p.ro16; e += kThreads) {
      const int c = e % p.ro16, d = e / p.ro16;
      Ys[d * p.ldy + c] = (c < ro && d < ri) ? A.xy_sign * Ri[d + c * pa] : 0.0;
    }
    ys_staged = true;
  }
  const int64_t ntiles = (A.rows + rpb - 1) / rpb;
  for (int64_t tile = bid; tile < ntiles; tile += nblk) {
    const int64_t row0 = tile * rpb;
    const int nrows = (int)min((int64_t)rpb, A.rows - row0);
    const int64_t i = row0 + lrow;
    const bool ok = lrow < nrows;
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = 0.0;
    if (ok) {
      if (A.parts) {
        // the partials were written by the previous launch, mostly into other
        // XCDs' L2: every load is a fabric round trip, so the loads of all 8
        // columns and of PC splits at a time are issued together (addresses
        // clamped into the array, the extra values discarded), then summed in
        // split order
        constexpr int PC = LEAN ? 2 : (sizeof(PT) == 4 ? 8 : 4) / (NRB >= 16 ? 2 : 1);  // (NRB 16 has no registers to spare)
        const int64_t sstride = (int64_t)A.rb * A.rows;
        const int cmax = ro - 1;
        for (int sp0 = 0; sp0 < A.nsplit; sp0 += PC) {
          PT w[8][PC];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int c = min(8 * cg + q, cmax);
#pragma unroll
            for (int u = 0; u < PC; ++u) {
              const int sp = min(sp0 + u, A.nsplit - 1);
              w[q][u] = A.parts[sp * sstride + (int64_t)c * A.rows + i];
            }
          }
#pragma unroll
          for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int u = 0; u < PC; ++u)
              if (sp0 + u < A.nsplit) v[q] += (double)w[q][u];
        }
        if (A.sum_out) {
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int c = 8 * cg + q;
            if (c < ro) A.sum_out[i + c * A.ld_sum] = v[q];
          }
        }
      }
      if (A.base) {
        if (A.st_d) {
          // neighbours of row i (with the periodic wrap when asked for)
          const int64_t n = A.rows;
          const bool hasdn = i + 1 < n || A.st_periodic;
          const bool hasup = i > 0 || A.st_periodic;
          const int64_t idn = i + 1 < n ? i + 1 : 0, iup = i > 0 ? i - 1 : n - 1;
          const double dd = A.st_d[i];
          const double edn = hasdn ? A.st_e[i] : 0.0, eup = hasup ? A.st_e[iup] : 0.0;
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int c = 8 * cg + q;
            if (c < ro) {
              const double* col = A.base + c * A.ld_base;
              double w = dd * col[i];
              if (hasdn) w += edn * col[idn];
              if (hasup) w += eup * col[iup];
              v[q] += A.base_sign * w;
            }
          }
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int c = 8 * cg + q;
            if (c < ro) v[q] += A.base_sign * A.base[i + c * A.ld_base];
          }
        }
      }
    }
    // the first XR of this thread's X values, loaded before Y is staged so that
    // both loads are in flight together
    constexpr int XR = LEAN ? 4 : 16;
    double xr[XR];
    if (hasX) {
#pragma unroll
      for (int u = 0; u < XR; ++u) {
        const int d = cg + u * tpr;
        xr[u] = (ok && d < ri) ? A.X[i + d * A.ldx] : 0.0;
      }
    }
    if (!ys_staged) {
      // xy_sign * op(Y) into LDS once, as Ys[d][c] (ri4 x ro16, zero padded);
      // the barrier below publishes it
      for (int e = t; e < p.ri4 * p.ro16; e += kThreads) {
        const int c = e % p.ro16, d = e / p.ro16;
        double y = 0.0;
        if (c < ro && d < ri) y = A.transY ? A.Y[c + d * A.ldy] : A.Y[d + c * A.ldy];
        Ys[d * p.ldy + c] = A.xy_sign * y;
      }
      ys_staged = true;
    }
    if (use_tile) {
      // the tile into LDS: partial outputs, X rows, W rows (zero beyond the data)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int c = 8 * cg + q;
        if (c < p.ro16) Ot[lrow * p.ldot + c] = (ok && c < ro) ? v[q] : 0.0;
      }
      for (int c = 8 * tpr + cg; c < p.ro16; c += tpr) Ot[lrow * p.ldot + c] = 0.0;
      if (hasX) {
#pragma unroll
        for (int u = 0; u < XR; ++u) {
          const int d = cg + u * tpr;
          if (d < p.ri4) Xt[lrow * p.ldxt + d] = xr[u];
        }
        for (int d = cg + XR * tpr; d < p.ri4; d += tpr) Xt[lrow * p.ldxt + d] = (ok && d < ri) ? A.X[i + d * A.ldx] : 0.0;
      }
      if (hasW && reduce)
        for (int c = cg; c < p.rw16; c += tpr) Wt[lrow * p.ldwt + c] = (ok && c < rw) ? A.W[i + c * A.ldw] : 0.0;
      __syncthreads();
      if (hasX) {
        // Ot += Xt * Ys, one 16 x 16 output block per wave at a time
        const int nrb = rpb / 16;
        for (int blk = wave; blk < p.nxy; blk += 4) {
          const int r0 = (blk % nrb) * 16, c0 = (blk / nrb) * 16;
          f64x4_t acc;
#pragma unroll
          for (int u = 0; u < 4; ++u) acc[u] = Ot[(r0 + 4 * u + lk) * p.ldot + c0 + lr];
          for (int d0 = 0; d0 < p.ri4; d0 += 4) {
            const double a = Xt[(r0 + lr) * p.ldxt + d0 + lk];
            const double b = Ys[(d0 + lk) * p.ldy + c0 + lr];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) Ot[(r0 + 4 * u + lk) * p.ldot + c0 + lr] = acc[u];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int c = 8 * cg + q;
          if (c < ro) v[q] = Ot[lrow * p.ldot + c];
        }
      }
    }
    if (ok) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int c = 8 * cg + q;
        if (c < ro) {
          if (A.out) A.out[i + c * A.ldo] = v[q];
          if (A.out2) A.out2[i + c * A.ldo2] = v[q];
          if (A.pk_f64) A.pk_f64[i * A.pk_cols + c] = v[q];
          if (A.pk_f32) A.pk_f32[pk_f32_offset(A.pk_layout, i, c, A.pk_cols)] = (float)v[q];
          if (A.pk_hi) {
            uint32_t hi, lo;
            split_bf16((float)v[q], hi, lo);
            const int64_t off = pk_offset(A.pk_layout, i, c, A.pk_cols);
            A.pk_hi[off] = (uint16_t)hi;
            A.pk_lo[off] = (uint16_t)lo;
          }
        }
      }
      // zero the packed columns beyond the current rank
      for (int c = ro + cg; c < A.pk_cols; c += tpr) {
        if (A.pk_f64) A.pk_f64[i * A.pk_cols + c] = 0.0;
        if (A.pk_f32) A.pk_f32[pk_f32_offset(A.pk_layout, i, c, A.pk_cols)] = 0.0f;
        if (A.pk_hi) {
          const int64_t off = pk_offset(A.pk_layout, i, c, A.pk_cols);
          A.pk_hi[off] = 0;
          A.pk_lo[off] = 0;
        }
      }
    }
    if (reduce) {
      // C += Ot' * Wt over the tile's rows (rows past the data are zero); with
      // gram_w the first rg16 block columns are Ot' * Ot
      const int nbi = p.ro16 / 16;
#pragma unroll
      for (int b = 0; b < NRB; ++b) {
        const int blk = wave + 4 * b;
        if (blk < p.nred) {
          const int c0 = (blk % nbi) * 16, j0 = (blk / nbi) * 16;
          const double* Bt = j0 < p.rg16 ? Ot + j0 : Wt + (j0 - p.rg16);
          const int ldb = j0 < p.rg16 ? p.ldot : ldwt;
          for (int k0 = 0; k0 < rpb; k0 += 4) {
            const double a = Ot[(k0 + lk) * p.ldot + c0 + lr];
            const double bb = Bt[(k0 + lk) * ldb + lr];
            racc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, racc[b], 0, 0, 0);
          }
        }
      }
    }
    if (use_tile) __syncthreads();  // the next tile overwrites the LDS tile
  }
  if (!reduce) return;
  double* gp = A.gpart + (int64_t)bid * ro * (rg + rw);
  const int nbi = p.ro16 / 16;
#pragma unroll
  for (int b = 0; b < NRB; ++b) {
    const int blk = wave + 4 * b;
    if (blk < p.nred) {
      const int c0 = (blk % nbi) * 16, j0 = (blk / nbi) * 16;
      const bool og = j0 < p.rg16;
      const int jl = (og ? j0 : j0 - p.rg16) + lr;  // column within its part
      const int cj = og ? jl : rg + jl;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int ci = c0 + 4 * u + lk;
        if (ci < ro && jl < (og ? ro : rw)) gp[ci + cj * ro] = racc[b][u];
      }
    }
  }
}

template <typename PT, int NRB, int CK = 0>
__global__ __launch_bounds__(kThreads) void k_rows(RowArgs<PT> A) {
  rows_body<PT, NRB, CK>(A, blockIdx.x, gridDim.x);
}

// The batched form: merged launches cover many narrow members with one tile per
// workgroup, so latency hiding matters more than loads in flight per thread:
// the lean body at two workgroups per CU for NRB <= 2 without a prologue.
template <typename PT, int NRB, int CK>
__global__ __launch_bounds__(kThreads, (NRB <= 2 && CK == 0) ? 2 : 1) void k_rows_batch(
    const unsigned char* __restrict__ tab, const int* __restrict__ start, int count) {
  const int b = blockIdx.x;
  const int i = batch_member(start, count, b);
  const RowArgs<PT>& A = *reinterpret_cast<const RowArgs<PT>*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes);
  rows_body<PT, NRB, CK, (NRB <= 2 && CK == 0)>(A, b - start[i], start[i + 1] - start[i]);
}

template <typename PT>
hipError_t launch_rows(RowArgs<PT> A, hipStream_t st) {
  const bool hasW = A.W != nullptr;
  const int rw = hasW ? A.rw : A.r_out;
  if (A.r_out < 1 || A.r_out > kMaxAug || rw < 1 || rw > kMaxAug || A.r_in < 0 || A.r_in > kMaxAug)
    return hipErrorInvalidValue;
  if (A.X && A.r_in < 1) return hipErrorInvalidValue;
  if (A.gram_w && (!hasW || !A.reduce)) return hipErrorInvalidValue;
  const RowsPlan p = rows_plan(A.r_out, A.r_in, rw, A.X != nullptr, hasW, A.reduce != 0, A.gram_w != 0);
  if (p.tpr > 16) return hipErrorInvalidValue;
  A.tr = p.rpb;
  const int rc = (A.gram_w ? A.r_out : 0) + rw;  // columns of C
  if (A.epi.stilde && (A.epi.kind != EPI_CHOL2 || !A.gram_w || rw != A.r_out)) return hipErrorInvalidValue;
  const bool pre = A.pre.kind != EPI_NONE;
  size_t pre_doubles = 0;
  if (pre) {
    if ((A.pre.kind != EPI_CHOL1 && A.pre.kind != EPI_CHOL2) || !A.pre_C || A.pre_a < 1 || A.pre_a > 32 ||
        A.pre_a != A.r_in || A.pre_a != A.r_out || !A.X || A.transY ||
        (A.pre_b != A.pre_a && A.pre_b != 2 * A.pre_a) || (A.pre.stilde && A.pre_b != 2 * A.pre_a))
      return hipErrorInvalidValue;
    pre_doubles = (size_t)A.pre_a * A.pre_b + epi_scratch_doubles(A.pre.kind, A.pre_a, A.pre_b);
  }
  const size_t lds = sizeof(double) * (p.lds_doubles + pre_doubles);
  const size_t epi_lds = sizeof(double) * ((size_t)A.r_out * rc + epi_scratch_doubles(A.epi.kind, A.r_out, rc));
  // room for the static shared variables (the Cholesky row buffers: 12.3 KB)
  if (lds > 147 * 1024 || epi_lds > 147 * 1024) return hipErrorInvalidValue;
  const int64_t ntiles = (A.rows + p.rpb - 1) / p.rpb;
  int grid = (int)(ntiles < A.max_blocks ? ntiles : A.max_blocks);
  if (grid < 1) grid = 1;
  const int nrb = (p.nred + 3) / 4;
  int nrbv = 0, ckv = 0;
  if (pre) {
    if (!std::is_same<PT, double>::value || nrb > 2) return hipErrorInvalidValue;
    ckv = A.pre_a <= 16 ? 16 : 32;
    nrbv = nrb <= 1 ? 1 : 2;
  } else {
    nrbv = nrb <= 1 ? 1 : (nrb <= 2 ? 2 : (nrb <= 4 ? 4 : (nrb <= 8 ? 8 : 16)));
  }
  const int kid = std::is_same<PT, double>::value ? KID_ROWS_D : KID_ROWS_F;
  if (!record_launch(kid, nrbv * 100 + ckv, (unsigned)grid, kThreads, lds, A)) {
    const dim3 g(grid), b(kThreads);
    switch (nrbv * 100 + ckv) {
      case 116: hipLaunchKernelGGL((k_rows<PT, 1, 16>), g, b, lds, st, A); break;
      case 216: hipLaunchKernelGGL((k_rows<PT, 2, 16>), g, b, lds, st, A); break;
      case 132: hipLaunchKernelGGL((k_rows<PT, 1, 32>), g, b, lds, st, A); break;
      case 232: hipLaunchKernelGGL((k_rows<PT, 2, 32>), g, b, lds, st, A); break;
      case 100: hipLaunchKernelGGL((k_rows<PT, 1>), g, b, lds, st, A); break;
      case 200: hipLaunchKernelGGL((k_rows<PT, 2>), g, b, lds, st, A); break;
      case 400: hipLaunchKernelGGL((k_rows<PT, 4>), g, b, lds, st, A); break;
      case 800: hipLaunchKernelGGL((k_rows<PT, 8>), g, b, lds, st, A); break;
      default: hipLaunchKernelGGL((k_rows<PT, 16>), g, b, lds, st, A); break;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !A.reduce) return e;
  const int ab = A.r_out * rc;
  double* Cr = A.epi.small->Cred;
  const ReduceLaunch R{(const double*)A.gpart, grid, ab, Cr};
  if (!record_launch(KID_REDUCE, 0, (unsigned)((ab + 63) / 64), kThreads, 0, R))
    hipLaunchKernelGGL(k_reduce, dim3((ab + 63) / 64), dim3(kThreads), 0, st, R.gpart, R.nb, R.ab, R.C);
  e = hipGetLastError();
  if (e != hipSuccess || A.epi.kind == EPI_NONE || A.defer_epi) return e;
  const bool chol = A.epi.kind == EPI_CHOL1 || A.epi.kind == EPI_CHOL2 || A.epi.kind == EPI_CHOL3;
  const int ck = (chol && A.r_out <= 16) ? 16 : ((chol && A.r_out <= 32) ? 32 : 64);
  const EpiLaunch L{A.epi, (const double*)Cr, A.r_out, rc};
  if (!record_launch(KID_EPILOGUE, ck, 1, kThreads, epi_lds, L)) {
    if (ck == 16)
      hipLaunchKernelGGL(k_epilogue<16>, dim3(1), dim3(kThreads), epi_lds, st, L.E, L.Cg, L.a, L.b);
    else if (ck == 32)
      hipLaunchKernelGGL(k_epilogue<32>, dim3(1), dim3(kThreads), epi_lds, st, L.E, L.Cg, L.a, L.b);
    else
      hipLaunchKernelGGL(k_epilogue<64>, dim3(1), dim3(kThreads), epi_lds, st, L.E, L.Cg, L.a, L.b);
  }
  return hipGetLastError();
}

template hipError_t launch_rows<double>(RowArgs<double>, hipStream_t);
template hipError_t launch_rows<float>(RowArgs<float>, hipStream_t);

// ----------------------------------------------------------------------------
// one-sided (Hestenes) Jacobi SVD of a k x k FP64 matrix, k <= 128, in one
// workgroup, then the truncation rank (rank_adaptive_unconventional.jl:222-223)
// r1 = min{rho >= 1 : sqrt(sum_{j >= rho} sigma_j^2) <= tol}, clipped to rmax.
// Pairs are processed in round-robin (circle-method) order, npairs at a time.
// ----------------------------------------------------------------------------

// Singular values (column norms of the rotated A, k x ke, ld lda), the sorted
// outputs U = A / sigma and V, and the truncation rank; nthr threads take part.
// pre: the sweeps ran on R' of S = Q R and Vm holds Q J, so U = Vm and V = A / sigma.
__device__ void jacobi_finish(const SvdArgs& P, const double* A, int lda, const double* Vm, int ldv, double* sig,
                              int nthr, bool pre = false) {
  const int k = P.k;
  const int ke = (k + 1) & ~1;
  const int t = threadIdx.x;
  for (int j = t; j < ke; j += nthr) {
    double s = 0.0;
    if (j < k)
      for (int i = 0; i < k; ++i) s += A[i + j * lda] * A[i + j * lda];
    sig[j] = (j < k) ? sqrt(s) : -1.0;
  }
  __syncthreads();
  for (int j = t; j < k; j += nthr) {
    int pos = 0;
    for (int l = 0; l < k; ++l) pos += (sig[l] > sig[j]) || (sig[l] == sig[j] && l < j);
    P.sigma[pos] = sig[j];
    const double inv = sig[j] > 0.0 ? 1.0 / sig[j] : 0.0;
    double* Ua = pre ? P.V : P.U;  // the normalised columns of A
    double* Uv = pre ? P.U : P.V;  // the accumulated rotations
    for (int i = 0; i < k; ++i) {
      Ua[i + pos * k] = A[i + j * lda] * inv;
      Uv[i + pos * k] = Vm[i + j * ldv];
    }
  }
  __syncthreads();
  if (t == 0) {
    // P.sigma holds the sorted values (written above by this workgroup)
    double tail = 0.0;
    int rho = k;
    for (int j = k - 1; j >= 1; --j) {
      double nt = tail + P.sigma[j] * P.sigma[j];
      if (sqrt(nt) <= P.tol) {
        tail = nt;
        rho = j;
      } else {
        break;
      }
    }
    if (rho > P.rmax) rho = (int)P.rmax;
    if (rho < 1) rho = 1;
    *P.r_out = rho;
  }
}

// x from lane (lane ^ m) of its row of 16 for the DPP patterns used below
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Sum over aligned groups of G lanes (G a power of two), left in every lane of
// the group.  Quad swaps (quad_perm [1,0,3,2] and [2,3,0,1]) make every quad
// uniform; the half-row mirror (lane i <- 7 - i) then pairs the two quads of
// an 8-lane group and the row mirror (i <- 15 - i) the two halves of a row.
template <int G>
__device__ __forceinline__ double group_sum(double x) {
  if (G >= 2) x += dpp_f64<0xB1>(x);
  if (G >= 4) x += dpp_f64<0x4E>(x);
  if (G >= 8) x += dpp_f64<0x141>(x);
  if (G >= 16) x += dpp_f64<0x140>(x);
  if (G >= 32) x += __shfl_xor(x, 16, 32);
  if (G >= 64) x += __shfl_xor(x, 32, 64);
  return x;
}

// group_sum for a group size known only at run time (uniform across the workgroup)
__device__ __forceinline__ double group_sum_rt(double x, int g) {
  switch (g) {
    case 1: return x;
    case 2: return group_sum<2>(x);
    case 4: return group_sum<4>(x);
    case 8: return group_sum<8>(x);
    case 16: return group_sum<16>(x);
    case 32: return group_sum<32>(x);
    default: return group_sum<64>(x);
  }
}

// QR preconditioning of the one-sided Jacobi SVD (Drmac and Veselic, "New
// fast and accurate Jacobi SVD algorithm", SIAM J. Matrix Anal. Appl. 29,
// 2008): with S = Q R, the sweeps run on R' instead of S.  Its columns are
// much closer to orthogonal than those of a graded S, so the sweeps to
// convergence drop from 15-30 to 6-11 at k = 16..64.  With R' J = W (J the
// accumulated rotations, W = Ur Sigma), S = (Q J) Sigma Ur': the left
// singular vectors are Q J and the right ones the normalised columns of W.
//
// Householder QR of A (k x k, ld lda), k <= 128, nt threads (a multiple of
// 64): reflector j (H_j = I - tau_j v v', v_j = 1, as LAPACK's dlarfg) in
// column j of H (k x k, ld k), tau_j in tau; on return A holds R' (lower
// triangular).  Wave 0 forms each reflector (rows lane and lane + 64);
// nt / 64 (k <= 64) or nt / 128 threads per column apply it.  A, H and tau may
// be in LDS or in global memory (the workgroup barriers order both).
__device__ void jacobi_qr_precondition(double* A, int lda, int k, double* H, double* tau, int nt) {
  __shared__ double s_hv[2];  // beta (the new diagonal), tau of the current reflector
  const int t = threadIdx.x;
  const int cap = k > 64 ? 128 : 64;
  const int gq = nt / cap;  // threads per column
  const int c = t / gq, ln = t % gq;
  for (int j = 0; j < k; ++j) {
    if (t < 64) {
      const int i0 = t, i1 = t + 64;
      const double x0 = i0 < k ? A[i0 + j * lda] : 0.0;
      const double x1 = i1 < k ? A[i1 + j * lda] : 0.0;
      const double a = j < 64 ? readlane_f64(x0, j) : readlane_f64(x1, j - 64);
      double s2 = (i0 > j && i0 < k ? x0 * x0 : 0.0) + (i1 > j && i1 < k ? x1 * x1 : 0.0);
      s2 = group_sum<64>(s2);
      double beta = a, tj = 0.0, scale = 0.0;
      if (s2 > 0.0) {
        beta = -copysign(sqrt(__builtin_fma(a, a, s2)), a);
        tj = (beta - a) / beta;
        scale = 1.0 / (a - beta);
      }
      if (i0 < k) H[i0 + j * k] = i0 == j ? 1.0 : (i0 > j ? x0 * scale : 0.0);
      if (i1 < k) H[i1 + j * k] = i1 == j ? 1.0 : (i1 > j ? x1 * scale : 0.0);
      if (t == 0) {
        s_hv[0] = beta;
        s_hv[1] = tj;
        tau[j] = tj;
      }
    }
    __syncthreads();
    const double tj = s_hv[1];
    if (c < k) {
      if (c > j && tj != 0.0) {
        double w = 0.0;
        for (int i = j + ln; i < k; i += gq) w = __builtin_fma(H[i + j * k], A[i + c * lda], w);
        w = group_sum_rt(w, gq) * tj;
        for (int i = j + ln; i < k; i += gq) A[i + c * lda] = __builtin_fma(-w, H[i + j * k], A[i + c * lda]);
      } else if (c == j) {
        for (int i = j + ln; i < k; i += gq) A[i + c * lda] = i == j ? s_hv[0] : 0.0;
      }
    }
    __syncthreads();
  }
  // A <- R' (each pair below/above the diagonal handled by one thread)
  for (int e = t; e < k * k; e += nt) {
    const int i = e % k, jj = e / k;
    if (i > jj) {
      A[i + jj * lda] = A[jj + i * lda];
      A[jj + i * lda] = 0.0;
    }
  }
  __syncthreads();
}

// X <- Q X = H_0 H_1 ... H_{k-1} X for the reflectors of jacobi_qr_precondition,
// X k rows x ncols (ld ldx, ncols <= 128), nt threads
__device__ void jacobi_apply_q(const double* H, const double* tau, int k, double* X, int ldx, int ncols, int nt) {
  const int t = threadIdx.x;
  const int cap = ncols > 64 ? 128 : 64;
  const int gq = nt / cap;
  const int c = t / gq, ln = t % gq;
  for (int j = k - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj != 0.0 && c < ncols) {
      double w = 0.0;
      for (int i = j + ln; i < k; i += gq) w = __builtin_fma(H[i + j * k], X[i + c * ldx], w);
      w = group_sum_rt(w, gq) * tj;
      for (int i = j + ln; i < k; i += gq) X[i + c * ldx] = __builtin_fma(-w, H[i + j * k], X[i + c * ldx]);
    }
    __syncthreads();
  }
}

constexpr int kJacobiWaveLda = 33;  // odd column stride against bank conflicts

// The same Jacobi sweeps for k <= 32 in one wavefront: G lanes per column pair,
// rows ln, ln + G, ... of the pair in each lane, the group sums through DPP,
// no workgroup barriers (LDS operations of one wavefront complete in order).
template <int G>
__device__ void jacobi_wave_body(const SvdArgs P, double* __restrict__ A, double* __restrict__ Vm,
                                 double* __restrict__ sig, double* __restrict__ Hq, double* __restrict__ tq) {
  constexpr int R = (128 / (G * G)) > 0 ? 128 / (G * G) : 1;  // rows per lane: ke <= 128 / G
  constexpr int LDA = kJacobiWaveLda;
  // rows i = ln + G j of every lane stay below 32 unless G = 64 (ke = 2)
  constexpr bool kGuard = G * R > 32;
  const int k = P.k;
  const int ke = (k + 1) & ~1;
  const int t = threadIdx.x;
  // the whole arrays: rows and columns past k (A) and ke (V) hold zeros, which
  // the rotations keep, so the row loops below need no bounds
  for (int e = t; e < LDA * 32; e += 64) {
    const int i = e % LDA, j = e / LDA;
    A[e] = (i < k && j < k) ? P.S[i + j * P.lds] : 0.0;
    Vm[e] = (i == j && i < ke) ? 1.0 : 0.0;
  }
  __syncthreads();
  jacobi_qr_precondition(A, LDA, k, Hq, tq, 64);  // A <- R' of S = Q R
  const int npairs = ke / 2;
  const int pair = t / G, ln = t % G;
  for (int sweep = 0; sweep < 60; ++sweep) {
    int rot = 0;
    for (int round = 0; round < ke - 1; ++round) {
      int pa = 0, pb = 0;
      const bool act = pair < npairs;
      if (pair == 0) {
        pa = 0;
        pb = 1 + (round % (ke - 1));
      } else if (act) {
        pa = 1 + ((round + pair) % (ke - 1));
        pb = 1 + ((round - pair + (ke - 1)) % (ke - 1));
      }
      double x[R], y[R], u[R], w[R];
      double al = 0.0, be = 0.0, ga = 0.0;
      // the pair's rows of A and of V, all loaded up front (one LDS wait)
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int i = ln + G * j;
        const bool ok = !kGuard || i < ke;  // inactive groups read column 0 and never rotate
        x[j] = ok ? A[i + pa * LDA] : 0.0;
        y[j] = ok ? A[i + pb * LDA] : 0.0;
        u[j] = ok ? Vm[i + pa * LDA] : 0.0;
        w[j] = ok ? Vm[i + pb * LDA] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < R; ++j) {
        al = __builtin_fma(x[j], x[j], al);
        be = __builtin_fma(y[j], y[j], be);
        ga = __builtin_fma(x[j], y[j], ga);
      }
      al = group_sum<G>(al);
      be = group_sum<G>(be);
      ga = group_sum<G>(ga);
      // rotate unless |ga| <= 4 eps sqrt(al be); tan of the angle
      // t = sign(d) 2ga / (|d| + sqrt(d^2 + 4ga^2)), d = be - al (|t| <= 1),
      // through refined hardware reciprocals (column norms are normal numbers)
      if (act && ga * ga > 16.0 * kEps * kEps * al * be) {
        const double d = be - al;
        const double g2 = 2.0 * ga;
        const double q = __builtin_fma(d, d, g2 * g2);
        const double den = fabs(d) + q * rsqrt_nr(q);
        double tt = g2 * rcp_nr(den);
        if (d < 0.0) tt = -tt;
        const double c = rsqrt_nr(__builtin_fma(tt, tt, 1.0));
        const double sn = c * tt;
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const int i = ln + G * j;
          if (!kGuard || i < ke) {
            A[i + pa * LDA] = c * x[j] - sn * y[j];
            A[i + pb * LDA] = sn * x[j] + c * y[j];
            Vm[i + pa * LDA] = c * u[j] - sn * w[j];
            Vm[i + pb * LDA] = sn * u[j] + c * w[j];
          }
        }
        rot = 1;
      }
      __builtin_amdgcn_wave_barrier();
    }
    if (!__any(rot)) break;
  }
  __syncthreads();
  jacobi_apply_q(Hq, tq, k, Vm, LDA, ke, 64);  // Vm <- Q J
  jacobi_finish(P, A, LDA, Vm, LDA, sig, 64, true);
}


// The workgroup form (k <= 128, blockDim.x threads): gsz lanes per column pair,
// a workgroup barrier per round.  A (k x ke) in LDS at an odd stride; V in
// LDS when it fits, else in P.work.
__device__ void jacobi_svd_body(const SvdArgs P, double* lds) {
  const int k = P.k;
  const int ke = (k + 1) & ~1;
  const int nt = blockDim.x, t = threadIdx.x;
  const int lda = k | 1;
  const int ldv = P.work ? ke : (ke | 1);
  double* A = lds;                      // k rows x ke columns, ld lda
  double* sig = A + (size_t)lda * ke;   // ke
  double* Vm = P.work ? P.work : sig + ke;  // ke x ke, ld ldv
  // the QR preconditioner's reflectors (k x k) and tau (k): after V, in LDS or
  // in the global work buffer
  double* Hq = Vm + (size_t)ldv * ke;
  double* tq = Hq + (size_t)k * k;
  for (int e = t; e < ke * k; e += nt) {
    const int i = e % k, j = e / k;
    A[i + j * lda] = (j < k) ? P.S[i + j * P.lds] : 0.0;
  }
  for (int e = t; e < ke * ke; e += nt) {
    const int i = e % ke, j = e / ke;
    Vm[i + j * ldv] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  jacobi_qr_precondition(A, lda, k, Hq, tq, nt);  // A <- R' of S = Q R
  const int npairs = ke / 2;
  int gsz = 1;
  while (gsz * 2 * npairs <= nt && gsz < 64) gsz *= 2;
  __shared__ int s_rot;
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (t == 0) s_rot = 0;
    __syncthreads();
    for (int round = 0; round < ke - 1; ++round) {
      const int pair = t / gsz, ln = t % gsz;
      int pa = -1, pb = -1;
      if (pair < npairs) {
        if (pair == 0) {
          pa = 0;
          pb = 1 + (round % (ke - 1));
        } else {
          pa = 1 + ((round + pair) % (ke - 1));
          pb = 1 + ((round - pair + (ke - 1)) % (ke - 1));
        }
      }
      double al = 0.0, be = 0.0, ga = 0.0;
      if (pa >= 0) {
        for (int i = ln; i < k; i += gsz) {
          const double x = A[i + pa * lda], y = A[i + pb * lda];
          al = __builtin_fma(x, x, al);
          be = __builtin_fma(y, y, be);
          ga = __builtin_fma(x, y, ga);
        }
      }
      al = group_sum_rt(al, gsz);
      be = group_sum_rt(be, gsz);
      ga = group_sum_rt(ga, gsz);
      // rotate unless |ga| <= 4 eps sqrt(al be) (as in the wavefront form)
      if (pa >= 0 && ga * ga > 16.0 * kEps * kEps * al * be) {
        const double d = be - al;
        const double g2 = 2.0 * ga;
        const double q = __builtin_fma(d, d, g2 * g2);
        const double den = fabs(d) + q * rsqrt_nr(q);
        double tt = g2 * rcp_nr(den);
        if (d < 0.0) tt = -tt;
        const double c = rsqrt_nr(__builtin_fma(tt, tt, 1.0));
        const double sn = c * tt;
        for (int i = ln; i < k; i += gsz) {
          const double x = A[i + pa * lda], y = A[i + pb * lda];
          A[i + pa * lda] = c * x - sn * y;
          A[i + pb * lda] = sn * x + c * y;
        }
        for (int i = ln; i < ke; i += gsz) {
          const double x = Vm[i + pa * ldv], y = Vm[i + pb * ldv];
          Vm[i + pa * ldv] = c * x - sn * y;
          Vm[i + pb * ldv] = sn * x + c * y;
        }
        if (ln == 0) s_rot = 1;
      }
      __syncthreads();
    }
    const int any = s_rot;
    __syncthreads();
    if (!any) break;
  }
  jacobi_apply_q(Hq, tq, k, Vm, ldv, ke, nt);  // Vm <- Q J
  jacobi_finish(P, A, lda, Vm, ldv, sig, nt, true);
}

__global__ __launch_bounds__(kThreads) void k_small_abt(const double* __restrict__ A, const double* __restrict__ B,
                                                        int k, double* __restrict__ C) {
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
    const int i = e % k, j = e / k;
    double s = 0.0;
    for (int l = 0; l < k; ++l) s += A[i + l * k] * B[j + l * k];
    C[e] = s;
  }
}

hipError_t launch_small_abt(const double* A, const double* B, int k, double* C, hipStream_t st) {
  if (k < 1 || k > kMaxAug) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_small_abt, dim3(1), dim3(kThreads), 0, st, A, B, k, C);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// r x r parts of the BUG adjoint
// ---------------------------------------------------------------------------

// X1 = R1^{-T} and X2 = R2^{-T} for upper triangular R1, R2 (r x r, ld r, in
// global memory): both are staged in LDS (ws: inv_ws_doubles(r)), thread j of
// wave 0 (R1) or wave 1 (R2) solves R x = e_j by back substitution into column
// j of an LDS block (ld r + 1 against bank conflicts), and the columns of
// R^{-1} are stored as the rows of X.  Called by all threads (>= 128).
__host__ __device__ inline int inv_ws_doubles(int r) { return 2 * r * r + 2 * r * (r + 1); }

__device__ void inv_transpose_upper2(const double* R1, double* X1, const double* R2, double* X2, int r,
                                     double* ws) {
  const int t = threadIdx.x, rr = r * r, lx = r + 1;
  double* R1s = ws;
  double* R2s = ws + rr;
  double* X1s = ws + 2 * rr;
  double* X2s = X1s + r * lx;
  for (int e = t; e < rr; e += blockDim.x) {
    R1s[e] = R1[e];
    R2s[e] = R2[e];
  }
  __syncthreads();
  const int which = t >> 6, j = t & 63;
  if (which < 2 && j < r) {
    const double* Rs = which ? R2s : R1s;
    double* x = (which ? X2s : X1s) + j * lx;
    for (int i = r - 1; i >= 0; --i) {
      double v = (i == j) ? 1.0 : 0.0;
      for (int l = i + 1; l <= j; ++l) v -= Rs[i + l * r] * x[l];  // x[l] = 0 for l > j
      x[i] = (i > j) ? 0.0 : v / Rs[i + i * r];
    }
  }
  __syncthreads();
  for (int e = t; e < rr; e += blockDim.x) {
    const int i = e % r, jj = e / r;  // (R^{-1})_{i jj} -> (R^{-T})_{jj i}
    X1[jj + i * r] = X1s[i + jj * lx];
    X2[jj + i * r] = X2s[i + jj * lx];
  }
}

// LDS of k_bug_adj_pre: r <= 32 stages Sb1, N, M, S0, T1, T2 (6 r^2) and runs
// the products on the matrix cores; larger r keeps T1, T2 only (2 r^2).  Then
// the inverses' workspace reuses it.
__host__ __device__ inline int bug_adj_pre_ws(int r) {
  const int prod = r <= 32 ? 6 * r * r : 2 * r * r;
  return prod > inv_ws_doubles(r) ? prod : inv_ws_doubles(r);
}

__global__ __launch_bounds__(kThreads) void k_bug_adj_pre(BugAdjSmall* A, const double* M, const double* N,
                                                          const double* S0, int lds, int r) {
  extern __shared__ double ws[];
  const int t = threadIdx.x, rr = r * r;
  if (r <= 32) {
    double* sSb = ws;
    double* sN = ws + rr;
    double* sM = ws + 2 * rr;
    double* sS = ws + 3 * rr;
    double* T1 = ws + 4 * rr;
    double* T2 = ws + 5 * rr;
    for (int e = t; e < rr; e += blockDim.x) {
      sSb[e] = A->Sb1[e];
      sN[e] = N[e];
      sM[e] = M[e];
      sS[e] = S0[e % r + (e / r) * lds];
    }
    __syncthreads();
    // T1 = Sb1 N, T2 = M S0
    wave_gemm(
        r, r, r, 2, [&](int b, int i, int l) { return b == 0 ? sSb[i + l * r] : sM[i + l * r]; },
        [&](int b, int l, int jj) { return b == 0 ? sN[l + jj * r] : sS[l + jj * r]; },
        [&](int b, int i, int jj, double v) { (b == 0 ? T1 : T2)[i + jj * r] = v; });
    __syncthreads();
    // Mb = T1 S0', Nb = Sb1' T2, Sb0 = M' T1
    wave_gemm(
        r, r, r, 3,
        [&](int b, int i, int l) { return b == 0 ? T1[i + l * r] : (b == 1 ? sSb[l + i * r] : sM[l + i * r]); },
        [&](int b, int l, int jj) { return b == 0 ? sS[jj + l * r] : (b == 1 ? T2[l + jj * r] : T1[l + jj * r]); },
        [&](int b, int i, int jj, double v) { (b == 0 ? A->Mb : (b == 1 ? A->Nb : A->Sb0))[i + jj * r] = v; });
  } else {
    double* T1 = ws;
    double* T2 = ws + rr;
    for (int e = t; e < rr; e += blockDim.x) {  // T1 = Sb1 N, T2 = M S0
      const int i = e % r, j = e / r;
      double a = 0.0, b = 0.0;
      for (int l = 0; l < r; ++l) {
        a += A->Sb1[i + l * r] * N[l + j * r];
        b += M[i + l * r] * S0[l + j * lds];
      }
      T1[e] = a;
      T2[e] = b;
    }
    __syncthreads();
    for (int e = t; e < rr; e += blockDim.x) {
      const int i = e % r, j = e / r;
      double mb = 0.0, nb = 0.0, sb = 0.0;
      for (int l = 0; l < r; ++l) {
        mb += T1[i + l * r] * S0[j + l * lds];    // (Sb1 N) S0'
        nb += A->Sb1[l + i * r] * T2[l + j * r];  // Sb1' (M S0)
        sb += M[l + i * r] * T1[l + j * r];       // M' (Sb1 N)
      }
      A->Mb[e] = mb;
      A->Nb[e] = nb;
      A->Sb0[e] = sb;
    }
  }
  __syncthreads();
  inv_transpose_upper2(A->RK, A->RKinvT, A->RL, A->RLinvT, r, ws);
}

__global__ __launch_bounds__(kThreads) void k_bug_adj_post(BugAdjSmall* A, int r) {
  for (int e = threadIdx.x; e < r * r; e += blockDim.x) {
    const int i = e % r, j = e / r;
    A->Sb0[e] += A->CK[j + i * r] + A->CL[e];
  }
}

__global__ __launch_bounds__(kThreads) void k_ksl_adj_small(BugAdjSmall* A, int r, int stage) {
  extern __shared__ double ws[];
  const int rr = r * r;
  if (stage == 0) {
    inv_transpose_upper2(A->RK, A->RKinvT, A->RL, A->RLinvT, r, ws);
    for (int e = threadIdx.x; e < rr; e += blockDim.x) {  // Mb = RL Sb1
      const int i = e % r, j = e / r;
      double v = 0.0;
      for (int l = 0; l < r; ++l) v += A->RL[i + l * r] * A->Sb1[l + j * r];
      A->Mb[e] = v;
    }
  } else if (stage == 1) {
    for (int e = threadIdx.x; e < rr; e += blockDim.x) {  // Nb = RK CL'
      const int i = e % r, j = e / r;
      double v = 0.0;
      for (int l = 0; l < r; ++l) v += A->RK[i + l * r] * A->CL[j + l * r];
      A->Nb[e] = v;
    }
  } else {
    for (int e = threadIdx.x; e < rr; e += blockDim.x) A->Sb0[e] = A->CK[(e / r) + (e % r) * r];
  }
}

hipError_t launch_ksl_adj_small(BugAdjSmall* A, int r, int stage, hipStream_t st) {
  if (r < 1 || r > kMaxRank || stage < 0 || stage > 2) return hipErrorInvalidValue;
  const size_t ws = stage == 0 ? sizeof(double) * inv_ws_doubles(r) : 0;
  hipLaunchKernelGGL(k_ksl_adj_small, dim3(1), dim3(kThreads), ws, st, A, r, stage);
  return hipGetLastError();
}

// Reverse of the rank-adaptive step, small parts; one workgroup per stage (see lri_kernels.h).
__global__ __launch_bounds__(kThreads) void k_ra_adj_small(RaAdjSmall* A, const SmallState* st, const double* S0,
                                                           int lds, int r, int r1, int stage) {
  extern __shared__ double ws[];  // stage 0: the inverses' workspace
  const int t = threadIdx.x, nt = blockDim.x, k2 = 2 * r;
  const double* Us = st->Ush;
  const double* Vs = st->Vsh;
  const double* sig = st->sigma;
  const double* M = st->M;  // k2 x r, ld k2
  const double* N = st->N;
  if (stage == 0) {
    for (int e = t; e < k2 * r1; e += nt) {  // Jm = Ush' Ubs, Km = Vsh' Vbs
      const int i = e % k2, j = e / k2;
      double a = 0.0, b = 0.0;
      for (int l = 0; l < k2; ++l) {
        a += Us[l + i * k2] * A->Ubs[l + j * k2];
        b += Vs[l + i * k2] * A->Vbs[l + j * k2];
      }
      A->Jm[e] = a;
      A->Km[e] = b;
    }
    __syncthreads();
    inv_transpose_upper2(A->RU, A->RUinvT, A->RV, A->RVinvT, r, ws);
  } else if (stage == 1) {
    for (int e = t; e < k2 * k2; e += nt) {
      const int i = e % k2, j = e / k2;
      double v = 0.0;
      if (i != j && (i < r1 || j < r1)) {
        const double F = 1.0 / (sig[j] * sig[j] - sig[i] * sig[i]);
        const double jij = j < r1 ? A->Jm[i + j * k2] : 0.0, jji = i < r1 ? A->Jm[j + i * k2] : 0.0;
        const double kij = j < r1 ? A->Km[i + j * k2] : 0.0, kji = i < r1 ? A->Km[j + i * k2] : 0.0;
        v = F * (jij - jji) * sig[j] + sig[i] * F * (kij - kji);
      }
      if (i == j && i < r1) v = A->Sb1[i + i * r1];
      A->core[e] = v;
    }
  } else if (stage == 2) {
    for (int e = t; e < k2 * k2; e += nt) {  // Shb <- core Vsh' (temporarily)
      const int i = e % k2, j = e / k2;
      double v = 0.0;
      for (int l = 0; l < k2; ++l) v += A->core[i + l * k2] * Vs[j + l * k2];
      A->Shb[e] = v;
    }
  } else if (stage == 3) {
    for (int e = t; e < k2 * k2; e += nt) {  // core <- Ush (core Vsh')
      const int i = e % k2, j = e / k2;
      double v = 0.0;
      for (int l = 0; l < k2; ++l) v += Us[i + l * k2] * A->Shb[l + j * k2];
      A->core[e] = v;
    }
  } else {
    // Shb = the cotangent of Shat (in core); Jm <- S0 N' (r x k2), Km <- S0' M' (r x k2), ld r
    for (int e = t; e < k2 * k2; e += nt) A->Shb[e] = A->core[e];
    for (int e = t; e < r * k2; e += nt) {
      const int i = e % r, l = e / r;
      double sn = 0.0, sm = 0.0;
      for (int q = 0; q < r; ++q) {
        sn += S0[i + q * lds] * N[l + q * k2];
        sm += S0[q + i * lds] * M[l + q * k2];
      }
      A->Jm[e] = sn;
      A->Km[e] = sm;
    }
    __syncthreads();
    const double* Sb = A->Shb;
    for (int e = t; e < r * k2; e += nt) {  // G1 = (S0 N') Shb', G2 = (S0' M') Shb  (r x k2)
      const int i = e % r, j = e / r;
      double g1 = 0.0, g2 = 0.0;
      for (int l = 0; l < k2; ++l) {
        g1 += A->Jm[i + l * r] * Sb[j + l * k2];
        g2 += A->Km[i + l * r] * Sb[l + j * k2];
      }
      A->G1[e] = g1;
      A->G2[e] = g2;
    }
    for (int e = t; e < k2 * r; e += nt) {  // G3 = Shb (N S0'), G4 = Shb' (M S0), core <- Shb N  (k2 x r)
      const int i = e % k2, j = e / k2;
      double g3 = 0.0, g4 = 0.0, sn = 0.0;
      for (int l = 0; l < k2; ++l) {
        g3 += Sb[i + l * k2] * A->Jm[j + l * r];
        g4 += Sb[l + i * k2] * A->Km[j + l * r];
        sn += Sb[i + l * k2] * N[l + j * k2];
      }
      A->G3[e] = g3;
      A->G4[e] = g4;
      A->core[e] = sn;
    }
    __syncthreads();
    for (int e = t; e < r * r; e += nt) {  // Sb0 = M' (Shb N)
      const int i = e % r, j = e / r;
      double v = 0.0;
      for (int l = 0; l < k2; ++l) v += M[l + i * k2] * A->core[l + j * k2];
      A->Sb0[e] = v;
    }
  }
}

hipError_t launch_ra_adj_small(RaAdjSmall* A, const SmallState* st, const double* S0, int lds, int r, int r1,
                               int stage, hipStream_t s) {
  if (r < 1 || r > kMaxRank || r1 < 1 || r1 > 2 * r || r1 > kMaxRank || stage < 0 || stage > 4)
    return hipErrorInvalidValue;
  const size_t ws = stage == 0 ? sizeof(double) * inv_ws_doubles(r) : 0;
  hipLaunchKernelGGL(k_ra_adj_small, dim3(1), dim3(kThreads), ws, s, A, st, S0, lds, r, r1, stage);
  return hipGetLastError();
}

hipError_t launch_bug_adj_pre(BugAdjSmall* A, const double* M, const double* N, const double* S0, int lds, int r,
                              hipStream_t st) {
  if (r < 1 || r > kMaxRank) return hipErrorInvalidValue;
  const size_t ws = sizeof(double) * (size_t)bug_adj_pre_ws(r);
  hipLaunchKernelGGL(k_bug_adj_pre, dim3(1), dim3(kThreads), ws, st, A, M, N, S0, lds, r);
  return hipGetLastError();
}

hipError_t launch_bug_adj_post(BugAdjSmall* A, int r, hipStream_t st) {
  if (r < 1 || r > kMaxRank) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_bug_adj_post, dim3(1), dim3(kThreads), 0, st, A, r);
  return hipGetLastError();
}

// workgroup 0 decomposes P0, workgroup 1 (if launched) P1
__global__ __launch_bounds__(1024) void k_jacobi_svd(SvdArgs P0, SvdArgs P1) {
  extern __shared__ double lds[];
  jacobi_svd_body(blockIdx.x == 0 ? P0 : P1, lds);
}

__device__ void jacobi_wave_dispatch(const SvdArgs P);

// one wavefront per decomposition, k <= 32 (workgroup b decomposes P_b)
__global__ __launch_bounds__(64) void k_jacobi_wave(SvdArgs P0, SvdArgs P1) {
  jacobi_wave_dispatch(blockIdx.x == 0 ? P0 : P1);
}

// the batched forms: one record holds the one or two decompositions of a launch
struct SvdPair {
  SvdArgs P0, P1;
};
__global__ __launch_bounds__(1024) void k_jacobi_svd_batch(const unsigned char* __restrict__ tab,
                                                           const int* __restrict__ start, int count) {
  extern __shared__ double lds[];
  const int i = batch_member(start, count, blockIdx.x);
  const SvdPair& S = *reinterpret_cast<const SvdPair*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes);
  jacobi_svd_body(blockIdx.x - start[i] == 0 ? S.P0 : S.P1, lds);
}
__global__ __launch_bounds__(64) void k_jacobi_wave_batch(const unsigned char* __restrict__ tab,
                                                          const int* __restrict__ start, int count) {
  const int i = batch_member(start, count, blockIdx.x);
  const SvdPair& S = *reinterpret_cast<const SvdPair*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes);
  jacobi_wave_dispatch(blockIdx.x - start[i] == 0 ? S.P0 : S.P1);
}

__device__ void jacobi_wave_dispatch(const SvdArgs P) {
  // one set of LDS arrays for every variant (25.7 KB: several decompositions per CU)
  __shared__ double A[kJacobiWaveLda * 32];
  __shared__ double Vm[kJacobiWaveLda * 32];
  __shared__ double sig[32];
  __shared__ double Hq[32 * 32];  // the QR preconditioner's reflectors
  __shared__ double tq[32];
  const int np = ((P.k + 1) & ~1) / 2;
  if (np > 8)
    jacobi_wave_body<4>(P, A, Vm, sig, Hq, tq);
  else if (np > 4)
    jacobi_wave_body<8>(P, A, Vm, sig, Hq, tq);
  else if (np > 2)
    jacobi_wave_body<16>(P, A, Vm, sig, Hq, tq);
  else if (np > 1)
    jacobi_wave_body<32>(P, A, Vm, sig, Hq, tq);
  else
    jacobi_wave_body<64>(P, A, Vm, sig, Hq, tq);
}

// LDS doubles of one decomposition; clears P.work when the right factor fits in LDS
static hipError_t jacobi_plan(SvdArgs& P, size_t* lds) {
  if (P.k < 1 || P.k > kMaxAug) return hipErrorInvalidValue;
  const int ke = (P.k + 1) & ~1;
  *lds = sizeof(double) * ((size_t)(P.k | 1) * ke + ke);  // A and sigma
  if (P.k <= 64) {
    // V, the QR preconditioner's reflectors and tau in LDS as well
    *lds += sizeof(double) * ((size_t)(ke | 1) * ke + (size_t)P.k * P.k + P.k);
    P.work = nullptr;
  } else if (!P.work) {
    return hipErrorInvalidValue;  // V, reflectors and tau in the global work buffer (kSvdWorkDoubles)
  }
  return hipSuccess;
}

// threads of the workgroup form: 8 or more lanes per column pair up to 1024
static int jacobi_threads(int k) { return k <= 64 ? 512 : 1024; }

hipError_t launch_jacobi_svd(SvdArgs P, hipStream_t st) {
  size_t lds;
  hipError_t e = jacobi_plan(P, &lds);
  if (e != hipSuccess) return e;
  const SvdPair S{P, P};
  if (P.k <= 32) {
    if (!record_launch(KID_JACOBI_WAVE, 0, 1, 64, 0, S))
      hipLaunchKernelGGL(k_jacobi_wave, dim3(1), dim3(64), 0, st, P, P);
  } else if (!record_launch(KID_JACOBI_WG, 0, 1, (unsigned)jacobi_threads(P.k), lds, S)) {
    hipLaunchKernelGGL(k_jacobi_svd, dim3(1), dim3(jacobi_threads(P.k)), lds, st, P, P);
  }
  return hipGetLastError();
}

hipError_t launch_jacobi_svd2(SvdArgs P0, SvdArgs P1, hipStream_t st) {
  size_t l0, l1;
  hipError_t e = jacobi_plan(P0, &l0);
  if (e == hipSuccess) e = jacobi_plan(P1, &l1);
  if (e != hipSuccess) return e;
  const SvdPair S{P0, P1};
  if (P0.k <= 32 && P1.k <= 32) {
    if (!record_launch(KID_JACOBI_WAVE, 0, 2, 64, 0, S))
      hipLaunchKernelGGL(k_jacobi_wave, dim3(2), dim3(64), 0, st, P0, P1);
  } else {
    const unsigned th = (unsigned)jacobi_threads(std::max(P0.k, P1.k));
    const size_t lds = l0 > l1 ? l0 : l1;
    if (!record_launch(KID_JACOBI_WG, 0, 2, th, lds, S))
      hipLaunchKernelGGL(k_jacobi_svd, dim3(2), dim3(th), lds, st, P0, P1);
  }
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// orthonormal completion of dropped CholeskyQR columns (see CompleteArgs)
// ----------------------------------------------------------------------------
constexpr int kCompleteThreads = 256;

// per thread v[0..7]: the workgroup sums into out[0..7] (LDS), visible to all after return
__device__ __forceinline__ void complete_block_sum8(double (&v)[8], double* red, double* out) {
#pragma unroll
  for (int q = 0; q < 8; ++q)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < 8; ++q) red[(threadIdx.x >> 6) * 8 + q] = v[q];
  __syncthreads();
  if (threadIdx.x < 8) {
    const int q = threadIdx.x;
    out[q] = (red[q] + red[8 + q]) + (red[16 + q] + red[24 + q]);
  }
  __syncthreads();
}

// h[0..n) = Q(:, 0..n)' y over the rows (h in LDS)
__device__ void complete_project(const double* __restrict__ Q, int64_t rows, int n, const double* __restrict__ y,
                                 double* red, double* h) {
  const int t = threadIdx.x;
  for (int c0 = 0; c0 < n; c0 += 8) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = t; i < rows; i += kCompleteThreads) {
      const double yi = y[i];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (c0 + q < n) v[q] = fma(Q[i + (c0 + q) * rows], yi, v[q]);
    }
    complete_block_sum8(v, red, h + c0);
  }
}

__global__ __launch_bounds__(kCompleteThreads) void k_complete_columns(CompleteArgs A) {
  __shared__ double s_R[kMaxRank * kMaxRank];  // R (k x k, ld k)
  __shared__ double s_h[kMaxAug + 8];
  __shared__ double s_red[32];
  __shared__ int s_dep[kMaxAug];
  __shared__ double s_min[4];
  __shared__ long long s_arg[4];
  const int t = threadIdx.x;
  const int k = A.k;
  const int64_t rows = A.rows;
  double* __restrict__ Q = A.Q;
  // one diagonal entry per thread: the common case (nothing dropped) costs one load
  if (!__syncthreads_or(t < k && (A.Rinv1[t + t * k] == 0.0 || A.Rinv2[t + t * k] == 0.0))) return;
  for (int e = t; e < k * k; e += kCompleteThreads) s_R[e] = 0.0;
  // the largest column norm of X, for the rounding-level threshold
  double cmax = 0.0;
  for (int c0 = 0; c0 < k; c0 += 8) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = t; i < rows; i += kCompleteThreads)
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (c0 + q < k) v[q] = fma(A.X[i + (c0 + q) * rows], A.X[i + (c0 + q) * rows], v[q]);
    complete_block_sum8(v, s_red, s_h);
    for (int q = 0; q < 8 && c0 + q < k; ++q) cmax = fmax(cmax, s_h[q]);
  }
  const double tol = 16.0 * k * kEps * sqrt(cmax);
  // classical Gram-Schmidt, each column projected out twice
  for (int j = 0; j < k; ++j) {
    double* __restrict__ qj = Q + (int64_t)j * rows;
    for (int64_t i = t; i < rows; i += kCompleteThreads) qj[i] = A.X[i + (int64_t)j * rows];
    __syncthreads();
    for (int pass = 0; pass < 2 && j > 0; ++pass) {
      complete_project(Q, rows, j, qj, s_red, s_h);
      for (int64_t i = t; i < rows; i += kCompleteThreads) {
        double v = qj[i];
        for (int c = 0; c < j; ++c) v = fma(-Q[i + c * rows], s_h[c], v);
        qj[i] = v;
      }
      if (t < j) s_R[t + j * k] += s_h[t];
      __syncthreads();
    }
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = t; i < rows; i += kCompleteThreads) v[0] = fma(qj[i], qj[i], v[0]);
    complete_block_sum8(v, s_red, s_h);
    const double nrm = sqrt(s_h[0]);
    const bool dep = !(nrm > tol);
    const double sc = dep ? 0.0 : 1.0 / nrm;
    for (int64_t i = t; i < rows; i += kCompleteThreads) qj[i] *= sc;
    if (t == 0) {
      s_R[j + j * k] = dep ? 0.0 : nrm;
      s_dep[j] = dep ? 1 : 0;
    }
    __syncthreads();
  }
  // orthonormal completion of the dependent columns (their rows of R are zero)
  for (int p = 0; p < k; ++p) {
    if (!s_dep[p]) continue;
    double* __restrict__ q = Q + (int64_t)p * rows;
    double best = 1e300;
    long long bi = 0;
    for (int64_t i = t; i < rows; i += kCompleteThreads) {
      double s2 = 0.0;
      for (int c = 0; c < k; ++c) s2 = fma(Q[i + c * rows], Q[i + c * rows], s2);
      if (s2 < best) {
        best = s2;
        bi = i;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off);
      const long long oi = __shfl_xor(bi, off);
      if (ob < best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if ((t & 63) == 0) {
      s_min[t >> 6] = best;
      s_arg[t >> 6] = bi;
    }
    __syncthreads();
    if (t == 0)
      for (int w = 1; w < 4; ++w)
        if (s_min[w] < s_min[0] || (s_min[w] == s_min[0] && s_arg[w] < s_arg[0])) {
          s_min[0] = s_min[w];
          s_arg[0] = s_arg[w];
        }
    __syncthreads();
    const int64_t is = s_arg[0];
    if (t < k) s_h[t] = Q[is + t * rows];
    __syncthreads();
    for (int64_t i = t; i < rows; i += kCompleteThreads) {  // column p is zero and takes no part
      double x = i == is ? 1.0 : 0.0;
      for (int c = 0; c < k; ++c) x = fma(-Q[i + c * rows], s_h[c], x);
      q[i] = x;
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      complete_project(Q, rows, k, q, s_red, s_h);
      if (t == 0) s_h[p] = 0.0;
      __syncthreads();
      for (int64_t i = t; i < rows; i += kCompleteThreads) {
        double x = q[i];
        for (int c = 0; c < k; ++c) x = fma(-Q[i + c * rows], s_h[c], x);
        q[i] = x;
      }
      __syncthreads();
    }
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = t; i < rows; i += kCompleteThreads) v[0] = fma(q[i], q[i], v[0]);
    complete_block_sum8(v, s_red, s_h);
    const double inv = 1.0 / sqrt(s_h[0]);
    for (int64_t i = t; i < rows; i += kCompleteThreads) q[i] *= inv;
    __syncthreads();
  }
  // the packed copy of Q for the next pass
  for (int64_t i = t; i < rows; i += kCompleteThreads)
    for (int p = 0; p < k; ++p) {
      const double x = Q[i + (int64_t)p * rows];
      if (A.pk_f64) A.pk_f64[i * A.pk_cols + p] = x;
      if (A.pk_f32) A.pk_f32[pk_f32_offset(A.pk_layout, i, p, A.pk_cols)] = (float)x;
      if (A.pk_hi) {
        uint32_t hi, lo;
        split_bf16((float)x, hi, lo);
        const int64_t off = pk_offset(A.pk_layout, i, p, A.pk_cols);
        A.pk_hi[off] = (uint16_t)hi;
        A.pk_lo[off] = (uint16_t)lo;
      }
    }
  // the products of R: Z = R (zt: R'), M = R - Q' W (mt: its transpose)
  if (A.Z)
    for (int e = t; e < k * k; e += kCompleteThreads) {
      const int i = e % k, j = e / k;
      if (A.zt) A.Z[j + i * A.ldz] = s_R[e];
      else A.Z[i + j * A.ldz] = s_R[e];
    }
  if (A.M)
    for (int c = 0; c < k; ++c) {
      complete_project(Q, rows, k, A.W + (int64_t)c * A.ldw, s_red, s_h);
      if (t < k) {
        const double x = s_R[t + c * k] - s_h[t];
        if (A.mt) A.M[c + t * A.ldm] = x;
        else A.M[t + c * A.ldm] = x;
      }
      __syncthreads();
    }
}

hipError_t launch_complete_columns(const CompleteArgs& A, hipStream_t st) {
  if (!A.X || !A.Q || A.rows < 1 || A.k < 1 || A.k > kMaxRank || !A.Rinv1 || !A.Rinv2 || (A.M && !A.W))
    return hipErrorInvalidValue;
  if (A.k > A.rows) return hipErrorInvalidValue;  // no room for an orthonormal completion
  hipLaunchKernelGGL(k_complete_columns, dim3(1), dim3(kCompleteThreads), 0, st, A);
  return hipGetLastError();
}

hipError_t launch_rowops_batched(int kid, int variant, const void* tab, const int* start, int count, unsigned total,
                                 unsigned block, size_t lds, hipStream_t st) {
  const unsigned char* t = static_cast<const unsigned char*>(tab);
  const dim3 g(total), b(block);
  switch (kid) {
    case KID_ROWS_D:
    case KID_ROWS_F: {
#define LRI_ROWS_BATCH(PT)                                                                                       \
  switch (variant) {                                                                                             \
    case 100: hipLaunchKernelGGL((k_rows_batch<PT, 1, 0>), g, b, lds, st, t, start, count); break;              \
    case 200: hipLaunchKernelGGL((k_rows_batch<PT, 2, 0>), g, b, lds, st, t, start, count); break;              \
    case 400: hipLaunchKernelGGL((k_rows_batch<PT, 4, 0>), g, b, lds, st, t, start, count); break;              \
    case 800: hipLaunchKernelGGL((k_rows_batch<PT, 8, 0>), g, b, lds, st, t, start, count); break;              \
    case 1600: hipLaunchKernelGGL((k_rows_batch<PT, 16, 0>), g, b, lds, st, t, start, count); break;            \
    default: return hipErrorInvalidValue;                                                                        \
  }
      if (kid == KID_ROWS_F) {
        LRI_ROWS_BATCH(float)
      } else {
        switch (variant) {
          case 116: hipLaunchKernelGGL((k_rows_batch<double, 1, 16>), g, b, lds, st, t, start, count); break;
          case 216: hipLaunchKernelGGL((k_rows_batch<double, 2, 16>), g, b, lds, st, t, start, count); break;
          case 132: hipLaunchKernelGGL((k_rows_batch<double, 1, 32>), g, b, lds, st, t, start, count); break;
          case 232: hipLaunchKernelGGL((k_rows_batch<double, 2, 32>), g, b, lds, st, t, start, count); break;
          default: { LRI_ROWS_BATCH(double) }
        }
      }
#undef LRI_ROWS_BATCH
      break;
    }
    case KID_REDUCE: hipLaunchKernelGGL(k_reduce_batch, g, b, lds, st, t, start, count); break;
    case KID_EPILOGUE:
      if (variant == 16) hipLaunchKernelGGL(k_epilogue_batch<16>, g, b, lds, st, t, start, count);
      else if (variant == 32) hipLaunchKernelGGL(k_epilogue_batch<32>, g, b, lds, st, t, start, count);
      else hipLaunchKernelGGL(k_epilogue_batch<64>, g, b, lds, st, t, start, count);
      break;
    case KID_JACOBI_WAVE: hipLaunchKernelGGL(k_jacobi_wave_batch, g, b, lds, st, t, start, count); break;
    case KID_JACOBI_WG: hipLaunchKernelGGL(k_jacobi_svd_batch, g, b, lds, st, t, start, count); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace lri
