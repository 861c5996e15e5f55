// Reverse (adjoint) of the matrix-ODE sub-steps with the quadratic right-hand
// side of lri_quad.hip,
//
//     K' = f(K) = A K + K M + g hv - N(K),  N_ic = sum_ab (D K)_ia K_ib T_abc,
//     S' = sgn (Ga S + S Gb + gaT' hbT - P(S)),  P_ij = sum_efab S_ea S_fb W_efi T_abj,
//
// integrated by N classical RK4 steps.  For a stage input X and the cotangent
// kbar of the stage's value f(X):
//
//     J_f(X)' kbar = A kbar + kbar M' - (D' P + Q),
//         P_ia = sum_bc kbar_ic X_ib T_abc,  Q_ib = sum_ac kbar_ic (D X)_ia T_abc,
//     Tbar_abc -= sum_i kbar_ic (D X)_ia X_ib,  Mbar += X' kbar,  hvbar += g' kbar,
//
// and the RK4 step y+ = y + h/6 (k1 + 2 k2 + 2 k3 + k4), X1 = y, X2 = y + h/2 k1,
// X3 = y + h/2 k2, X4 = y + h k3, runs backwards stage by stage:
//     kbar_s = (h w_s / 6) ybar+ + c_s h Xbar_{s+1},  Xbar_s = J_f(X_s)' kbar_s,
//     ybar = ybar+ + sum_s Xbar_s.
// The K and L sub-problems need neighbouring rows (A and D are tridiagonal), so
// a stage is one launch over all rows; the r x r S sub-problem runs in one
// workgroup, and the basis cotangents from the Galerkin data (X'AX, F'X and the
// tensor sum_i (D X)_ia X_ib X_ic) are one row kernel.
#include <algorithm>

#include "lri_common.h"
#include "lri_kernels.h"

namespace lri {

namespace {

__device__ __forceinline__ double stage_c(int s) { return s == 2 ? 1.0 : 0.5; }
__device__ __forceinline__ double stage_w(int s) { return (s == 1 || s == 2) ? 2.0 : 1.0; }

// Rows r0 - 1 .. r0 + R of a rows x k matrix (ld) into buf[lr * k + c], a row
// outside [0, n) taking its periodic image (only read by periodic stencils) and
// rows past the image range reading as zero.
__device__ __forceinline__ void stage_rows(double* buf, const double* src, int64_t ld, int64_t r0, int R, int k,
                                           int64_t n) {
  for (int e = threadIdx.x; e < (R + 2) * k; e += blockDim.x) {
    const int lr = e % (R + 2), c = e / (R + 2);
    int64_t ir = r0 + lr - 1;
    if (ir < 0) ir += n;
    else if (ir >= n) ir -= n;
    buf[lr * k + c] = (ir >= 0 && ir < n) ? src[ir + c * ld] : 0.0;
  }
}

// (D x)_i for staged rows (x[lr * k + c] holds row r0 + lr - 1); D = I when dd == nullptr
__device__ __forceinline__ double apply_d(const double* x, int lr, int c, int k, int64_t i, int64_t n,
                                          const double* dl, const double* dd, const double* du, int dper) {
  const double x0 = x[(lr + 1) * k + c];
  if (!dd) return x0;
  const bool hasl = i > 0 || dper, hasu = i < n - 1 || dper;
  return dd[i] * x0 + (hasl ? dl[i] * x[lr * k + c] : 0.0) + (hasu ? du[i] * x[(lr + 2) * k + c] : 0.0);
}

// (D' y)_i: column i of D is dd_i (row i), du_{i-1} (row i - 1), dl_{i+1} (row i + 1)
__device__ __forceinline__ double apply_dt(const double* y, int lr, int c, int k, int64_t i, int64_t n,
                                           const double* dl, const double* dd, const double* du, int dper) {
  const double y0 = y[(lr + 1) * k + c];
  if (!dd) return y0;
  const bool hasl = i > 0 || dper, hasu = i < n - 1 || dper;
  const int64_t im = i > 0 ? i - 1 : n - 1, ip = i < n - 1 ? i + 1 : 0;
  return dd[i] * y0 + (hasl ? du[im] * y[lr * k + c] : 0.0) + (hasu ? dl[ip] * y[(lr + 2) * k + c] : 0.0);
}

// (A y)_i for the symmetric tridiagonal A (st_e[i] couples rows i and i + 1,
// st_e[n - 1] the last and first row when periodic)
__device__ __forceinline__ double apply_a(const double* y, int lr, int c, int k, int64_t i, int64_t n,
                                          const double* st_d, const double* st_e, int periodic) {
  double v = st_d[i] * y[(lr + 1) * k + c];
  if (i > 0) v += st_e[i - 1] * y[lr * k + c];
  else if (periodic) v += st_e[n - 1] * y[lr * k + c];
  if (i < n - 1) v += st_e[i] * y[(lr + 2) * k + c];
  else if (periodic) v += st_e[n - 1] * y[(lr + 2) * k + c];
  return v;
}

}  // namespace

// ----------------------------------------------------------------------------
// One reverse RK4 stage over all rows.  Workgroup: tiles of 16 rows (4 waves);
// a workgroup walks the tiles blockIdx.x, blockIdx.x + grid, ... and adds their
// parameter cotangents into its own slice of Q.part (no atomics: the sum over
// workgroups is taken later, in order).  The three tensor contractions of a
// tile are products on the FP64 matrix cores, T read in its storage order
// T[c + k (a + k b)]:
//   P (18 x k, the tile and its halo rows) = Z1 T1,  Z1_i,(c + k b) = X_ib kbar_ic,
//   Q (16 x k) = Z2 T2,  Z2_i,(c + k a) = (D X)_ia kbar_ic,
//   Tbar (k^2 x k) -= Z3' kbar,  Z3_i,(a + k b) = (D X)_ia X_ib,
// wave w taking column block w % nct (nct = ceil(k / 16)) and the w / nct-th
// slice of the k^2 reduction of P and Q (summed through LDS), and every fourth
// 16 x 16 block of Tbar.
// ----------------------------------------------------------------------------

constexpr int kAdjMaxGrid = 512;
constexpr int kAdjRows = 16;

int quad_adj_grid(int64_t rows, int k) {
  (void)k;
  const int64_t tiles = (rows + kAdjRows - 1) / kAdjRows;
  return (int)std::min<int64_t>(tiles, kAdjMaxGrid);
}

__global__ __launch_bounds__(256) void k_quad_adj_stage(QuadAdjStageArgs Q) {
  constexpr int R = kAdjRows;
  __shared__ double xs[(R + 2) * kQuadMaxRank];  // stage input rows r0 - 1 .. r0 + 16: [lr * k + c]
  __shared__ double ks[(R + 2) * kQuadMaxRank];  // kbar, same rows
  __shared__ double dxs[R * kQuadMaxRank];       // (D X), rows r0 .. r0 + 15 (zero past n)
  __shared__ double pP[4][(R + 2) * kQuadMaxRank];  // slices of P, [row * 32 + a]
  __shared__ double pQ[4][R * kQuadMaxRank];        // slices of Q
  __shared__ double red[256];
  const int k = Q.k, t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, lr = lane & 15, lk = lane >> 4;
  const int64_t n = Q.n;
  const int s = Q.stage;
  const double wy = Q.h * stage_w(s) / 6.0;
  const double wx = s < 3 ? stage_c(s) * Q.h : 0.0;
  const int k2 = k * k, k3 = k2 * k;
  const int plen = quad_adj_part_len(k, Q.q);
  double* part = Q.part + (int64_t)blockIdx.x * plen;
  const int64_t ntiles = (n + R - 1) / R;
  const int nct = (k + 15) / 16, nch = 4 / nct;
  const int ct = wave % nct, ch = wave / nct;
  const int clen = (k2 + 4 * nch - 1) / (4 * nch) * 4;
  const int kbeg = ch * clen, kend = min(kbeg + clen, k2);
  const int col = ct * 16 + lr;  // this lane's output column in P and Q
  double scale_acc = 0.0;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t r0 = tile * R;
    __syncthreads();  // the previous tile's LDS reads are done
    for (int e = t; e < (R + 2) * k; e += blockDim.x) {
      const int lrow = e % (R + 2), c = e / (R + 2);
      int64_t ir = r0 + lrow - 1;
      if (ir < 0) ir += n;
      else if (ir >= n) ir -= n;
      double x = 0.0, kb = 0.0;
      if (ir >= 0 && ir < n) {
        const int64_t o = ir + c * Q.ld;
        x = Q.X[o];
        kb = wy * Q.Ybar[o] + (s < 3 ? wx * Q.Xbp[o] : 0.0);
      }
      xs[lrow * k + c] = x;
      ks[lrow * k + c] = kb;
    }
    __syncthreads();
    if (Q.T) {
      for (int e = t; e < R * k; e += blockDim.x) {
        const int lrow = e / k, a = e % k;
        const int64_t i = r0 + lrow;
        dxs[e] = i < n ? apply_d(xs, lrow, a, k, i, n, Q.dl, Q.dd, Q.du, Q.dper) : 0.0;
      }
      // P on the staged rows 0..15 (acc1) and 16, 17 (acc2), same tensor operand
      f64x4_t acc1 = {0.0, 0.0, 0.0, 0.0}, acc2 = {0.0, 0.0, 0.0, 0.0};
      for (int kk0 = kbeg; kk0 < kend; kk0 += 4) {
        const int kk = kk0 + lk;  // (c, b) = (kk % k, kk / k)
        double a1 = 0.0, a2 = 0.0, b = 0.0;
        if (kk < kend) {
          const int c = kk % k, bb = kk / k;
          a1 = xs[lr * k + bb] * ks[lr * k + c];
          if (lr < 2) a2 = xs[(R + lr) * k + bb] * ks[(R + lr) * k + c];
          if (col < k) b = Q.T[c + (int64_t)k * (col + (int64_t)k * bb)];
        }
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b, acc1, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b, acc2, 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        pP[ch][(4 * u + lk) * kQuadMaxRank + col] = acc1[u];
        if (4 * u + lk < 2) pP[ch][(R + 4 * u + lk) * kQuadMaxRank + col] = acc2[u];
      }
    }
    __syncthreads();  // dxs ready
    if (Q.T) {
      f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
      for (int kk0 = kbeg; kk0 < kend; kk0 += 4) {
        const int kk = kk0 + lk;  // (c, a) = (kk % k, kk / k)
        double a = 0.0, b = 0.0;
        if (kk < kend) {
          a = dxs[lr * k + kk / k] * ks[(lr + 1) * k + kk % k];
          if (col < k) b = Q.T[kk + (int64_t)k2 * col];
        }
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) pQ[ch][(4 * u + lk) * kQuadMaxRank + col] = acc[u];
      // Tbar -= Z3' kbar over the tile's rows: 16 x 16 blocks (a + k b rows, c columns)
      const int nab = (k2 + 15) / 16;
      for (int blk = wave; blk < nab * nct; blk += 4) {
        const int ab0 = (blk % nab) * 16, c0 = (blk / nab) * 16;
        const int ab = ab0 + lr, c = c0 + lr;
        f64x4_t tacc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i0 = 0; i0 < R; i0 += 4) {
          const int i = i0 + lk;
          const double a = ab < k2 ? dxs[i * k + ab % k] * xs[(i + 1) * k + ab / k] : 0.0;
          const double b = c < k ? ks[(i + 1) * k + c] : 0.0;
          tacc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, tacc, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int abu = ab0 + 4 * u + lk;
          if (abu < k2 && c < k) part[c + (int64_t)k * abu] -= tacc[u];
        }
      }
    }
    __syncthreads();  // P and Q slices ready
    const int nr = (int)std::min<int64_t>(R, n - r0);
    for (int e = t; e < R * k; e += blockDim.x) {
      const int lrow = e % R, c = e / R;
      const int64_t i = r0 + lrow;
      if (i >= n) continue;
      const double* kr = ks + (lrow + 1) * k;
      double xb = 0.0;
      if (Q.st_d) {
        xb = apply_a(ks, lrow, c, k, i, n, Q.st_d, Q.st_e, Q.periodic);
        scale_acc += kr[c] * apply_a(xs, lrow, c, k, i, n, Q.st_d, Q.st_e, Q.periodic);
      }
      if (Q.M)
        for (int l = 0; l < k; ++l) xb += kr[l] * Q.M[c + l * k];
      if (Q.T) {
        // (D' P)_ic from the summed slices of P on rows i - 1, i, i + 1 (staged lrow, lrow + 1, lrow + 2)
        double pm = 0.0, p0 = 0.0, pp = 0.0, qv = 0.0;
        for (int h = 0; h < nch; ++h) {
          pm += pP[h][lrow * kQuadMaxRank + c];
          p0 += pP[h][(lrow + 1) * kQuadMaxRank + c];
          pp += pP[h][(lrow + 2) * kQuadMaxRank + c];
          qv += pQ[h][lrow * kQuadMaxRank + c];
        }
        double dp = p0;
        if (Q.dd) {
          const bool hasl = i > 0 || Q.dper, hasu = i < n - 1 || Q.dper;
          const int64_t im = i > 0 ? i - 1 : n - 1, ip = i < n - 1 ? i + 1 : 0;
          dp = Q.dd[i] * p0 + (hasl ? Q.du[im] * pm : 0.0) + (hasu ? Q.dl[ip] * pp : 0.0);
        }
        xb -= dp + qv;
      }
      const int64_t o = i + c * Q.ld;
      Q.Xbo[o] = xb;
      Q.Acc[o] = (s == 3 ? Q.Ybar[o] : Q.Acc[o]) + xb;
    }
    if (Q.M) {
      for (int e = t; e < k * k; e += blockDim.x) {  // Mbar_lc = sum_i X_il kbar_ic
        const int l = e % k, c = e / k;
        double v = 0.0;
        for (int lrow = 0; lrow < nr; ++lrow) v += xs[(lrow + 1) * k + l] * ks[(lrow + 1) * k + c];
        part[k3 + e] += v;
      }
    }
    for (int e = t; e < Q.q * k; e += blockDim.x) {  // hvbar_qc = sum_i g_iq kbar_ic
      const int qq = e % Q.q, c = e / Q.q;
      double v = 0.0;
      for (int lrow = 0; lrow < nr; ++lrow) v += Q.g[(r0 + lrow) + qq * Q.ldg] * ks[(lrow + 1) * k + c];
      part[k3 + k * k + e] += v;
    }
  }
  red[t] = scale_acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) part[plen - 1] += red[0];
}

hipError_t launch_quad_adj_stage(QuadAdjStageArgs Q, hipStream_t st) {
  if (Q.k < 1 || Q.k > kQuadMaxRank || Q.n < 1 || Q.stage < 0 || Q.stage > 3 || Q.q < 0 || Q.q > kMaxForce ||
      !Q.part || !Q.X || !Q.Ybar || !Q.Xbo || !Q.Acc || (Q.stage < 3 && !Q.Xbp) || (Q.q > 0 && !Q.g) ||
      (Q.st_d && !Q.st_e))
    return hipErrorInvalidValue;
  Q.grid = quad_adj_grid(Q.n, Q.k);
  hipLaunchKernelGGL(k_quad_adj_stage, dim3(Q.grid), dim3(256), 0, st, Q);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_part_sum(const double* __restrict__ part, int count, int len,
                                                  double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  double s = 0.0;
  for (int g = 0; g < count; ++g) s += part[(int64_t)g * len + e];
  out[e] = s;
}

hipError_t launch_part_sum(const double* part, int count, int len, double* out, hipStream_t st) {
  if (count < 1 || len < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_part_sum, dim3((len + 255) / 256), dim3(256), 0, st, part, count, len, out);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Reverse of the S sub-step, one workgroup of 1024 threads (16 waves).  The
// k x l matrices live in LDS; the third-order intermediates in three buffers
// B0, B1, B2 (LDS when they and T fit, else Q.work), reused as they die:
//   forward f:  Y_ebc = sum_f X_fb W_efc -> B0,  Z_abc = sum_e X_ea Y_ebc -> B1,
//               P = Z' T (k x l);
//   reverse:    Y -> B0, Z -> B1;  Tbar_abj -= sum_i Z_abi pb_ij, X1_ebj = sum_a X_ea T_abj -> B0;
//               Zt_efj = sum_b X_fb X1_ebj -> B1;  Wbar_efi -= sum_j pb_ij Zt_efj,
//               C_efj = sum_i pb_ij W_efi -> B0;  D1_ebj = sum_f C_efj X_fb -> B1,
//               E1_faj = sum_e C_efj X_ea -> B2;
//               J_P' pb = sum_bj D1_ibj T_abj + sum_a'j E1_ia'j T_a'aj,
// every contraction a product on the FP64 matrix cores (wave_gemm).
// ----------------------------------------------------------------------------

namespace {

struct SAdjBufs {
  double *B0, *B1, *B2;
  const double* T;  // the tensor T (LDS copy or Q.T)
};

// out = f(X) (k x l, ld k); tmp: k x l scratch
__device__ void sadj_f(const QuadSAdjArgs& Q, const double* X, double* out, const double* cb, double* tmp,
                       const SAdjBufs& w) {
  const int k = Q.k, l = Q.l, t = threadIdx.x, nt = blockDim.x;
  const double* W = Q.W;
  const double* T = w.T;
  double* B0 = w.B0;
  double* B1 = w.B1;
  if (W) {
    wave_gemm(k * k, l, k, 1, [&](int, int m, int kk) { return W[m + k * k * kk]; },
              [&](int, int kk, int nn) { return X[kk + nn * k]; },
              [&](int, int m, int nn, double v) { B0[m + k * k * nn] = v; });
    __syncthreads();
    wave_gemm(k, l, k, l, [&](int b, int c, int e) { return B0[c + k * e + k * k * b]; },
              [&](int, int e, int a) { return X[e + a * k]; },
              [&](int b, int c, int a, double v) { B1[c + k * a + k * l * b] = v; });
    __syncthreads();
    wave_gemm(k, l, l * l, 1, [&](int, int i, int ab) { return B1[i + k * ab]; },
              [&](int, int ab, int j) { return T[j + l * ab]; },
              [&](int, int i, int j, double v) { tmp[i + j * k] = v; });
    __syncthreads();
  }
  for (int e = t; e < k * l; e += nt) {
    const int i = e % k, j = e / k;
    double v = cb[e];
    for (int m = 0; m < k; ++m) v += Q.Ga[i + m * k] * X[m + j * k];
    for (int m = 0; m < l; ++m) v += X[i + m * k] * Q.Gb[m + j * l];
    if (W) v -= tmp[e];
    out[e] = Q.sgn * v;
  }
  __syncthreads();
}

// xb = J_f(X)' kb; accumulates Gab, Gbb, cbar (LDS) and Wb, Tb (global)
__device__ void sadj_rev(const QuadSAdjArgs& Q, const double* X, const double* kb, double* pb, double* xb,
                         double* tmp, double* Gab, double* Gbb, double* cbar, const SAdjBufs& w) {
  const int k = Q.k, l = Q.l, t = threadIdx.x, nt = blockDim.x;
  const int kk2 = k * k, ll2 = l * l;
  const double* W = Q.W;
  const double* T = w.T;
  double* B0 = w.B0;
  double* B1 = w.B1;
  double* B2 = w.B2;
  for (int e = t; e < k * l; e += nt) pb[e] = Q.sgn * kb[e];
  __syncthreads();
  if (W) {
    wave_gemm(kk2, l, k, 1, [&](int, int m, int kk) { return W[m + kk2 * kk]; },  // Y -> B0
              [&](int, int kk, int nn) { return X[kk + nn * k]; },
              [&](int, int m, int nn, double v) { B0[m + kk2 * nn] = v; });
    __syncthreads();
    wave_gemm(k, l, k, l, [&](int b, int c, int e) { return B0[c + k * e + kk2 * b]; },  // Z -> B1
              [&](int, int e, int a) { return X[e + a * k]; },
              [&](int b, int c, int a, double v) { B1[c + k * a + k * l * b] = v; });
    __syncthreads();
    wave_gemm(l, ll2, k, 1, [&](int, int j, int i) { return pb[i + j * k]; },  // Tbar -= pb' Z
              [&](int, int i, int ab) { return B1[i + k * ab]; },
              [&](int, int j, int ab, double v) { Q.Tb[j + l * ab] -= v; });
    wave_gemm(k, l, l, l, [&](int b, int e, int a) { return X[e + a * k]; },  // X1 -> B0
              [&](int b, int a, int j) { return T[j + l * a + ll2 * b]; },
              [&](int b, int e, int j, double v) { B0[e + k * b + k * l * j] = v; });
    __syncthreads();
    wave_gemm(k, k, l, l, [&](int j, int e, int b) { return B0[e + k * b + k * l * j]; },  // Zt -> B1
              [&](int, int b, int f) { return X[f + b * k]; },
              [&](int j, int e, int f, double v) { B1[e + k * f + kk2 * j] = v; });
    __syncthreads();
    wave_gemm(k, kk2, l, 1, [&](int, int c, int j) { return pb[c + j * k]; },  // Wbar -= pb Zt
              [&](int, int j, int ef) { return B1[ef + kk2 * j]; },
              [&](int, int c, int ef, double v) { Q.Wb[c + k * ef] -= v; });
    wave_gemm(kk2, l, k, 1, [&](int, int ef, int i) { return W[i + k * ef]; },  // C -> B0
              [&](int, int i, int j) { return pb[i + j * k]; },
              [&](int, int ef, int j, double v) { B0[ef + kk2 * j] = v; });
    __syncthreads();
    wave_gemm(k, l, k, l, [&](int j, int e, int f) { return B0[e + k * f + kk2 * j]; },  // D1 -> B1
              [&](int, int f, int b) { return X[f + b * k]; },
              [&](int j, int e, int b, double v) { B1[e + k * b + k * l * j] = v; });
    wave_gemm(k, l, k, l, [&](int j, int f, int e) { return B0[e + k * f + kk2 * j]; },  // E1 -> B2
              [&](int, int e, int a) { return X[e + a * k]; },
              [&](int j, int f, int a, double v) { B2[f + k * a + k * l * j] = v; });
    __syncthreads();
    wave_gemm(k, l, 2 * ll2, 1,  // J_P' pb -> tmp
              [&](int, int i, int q) { return q < ll2 ? B1[i + k * q] : B2[i + k * (q - ll2)]; },
              [&](int, int q, int a) {
                if (q < ll2) return T[q / l + l * a + ll2 * (q % l)];  // T_{a b j}, q = b + l j
                const int q2 = q - ll2;
                return T[q2 / l + l * (q2 % l) + ll2 * a];  // T_{b a j}
              },
              [&](int, int i, int a, double v) { tmp[i + a * k] = v; });
    __syncthreads();
  }
  for (int e = t; e < k * l; e += nt) {
    const int i = e % k, a = e / k;
    double v = 0.0;
    for (int m = 0; m < k; ++m) v += Q.Ga[m + i * k] * pb[m + a * k];
    for (int m = 0; m < l; ++m) v += pb[i + m * k] * Q.Gb[a + m * l];
    if (W) v -= tmp[e];
    xb[e] = v;
  }
  for (int e = t; e < k * k; e += nt) {  // Gab_im += sum_a pb_ia X_ma
    const int i = e % k, m = e / k;
    double v = 0.0;
    for (int a = 0; a < l; ++a) v += pb[i + a * k] * X[m + a * k];
    Gab[e] += v;
  }
  for (int e = t; e < l * l; e += nt) {  // Gbb_mj += sum_i X_im pb_ij
    const int m = e % l, j = e / l;
    double v = 0.0;
    for (int i = 0; i < k; ++i) v += X[i + m * k] * pb[i + j * k];
    Gbb[e] += v;
  }
  for (int e = t; e < k * l; e += nt) cbar[e] += pb[e];
  __syncthreads();
}

}  // namespace

// LDS doubles: 11 k l + k^2 + l^2 for the matrices, then (when they fit) the
// three intermediate buffers and T
__host__ __device__ inline int quad_sadj_small_doubles(int k, int l) { return 11 * k * l + k * k + l * l; }
__host__ __device__ inline int quad_sadj_r3(int k, int l) { return k * l * (k > l ? k : l); }
constexpr int kSAdjLdsMax = 20480;  // 160 KB

template <bool kLds>
__global__ __launch_bounds__(1024) void k_quad_sadj(QuadSAdjArgs Q) {
  extern __shared__ double lds[];
  const int k = Q.k, l = Q.l, kl = k * l, t = threadIdx.x, nt = blockDim.x;
  double* y = lds;
  double* fv = y + kl;
  double* x = fv + kl;
  double* kb = x + kl;
  double* pb = kb + kl;
  double* xb = pb + kl;
  double* yb = xb + kl;
  double* acc = yb + kl;
  double* cb = acc + kl;
  double* cbar = cb + kl;
  double* tmp = cbar + kl;
  double* Gab = tmp + kl;
  double* Gbb = Gab + k * k;
  const int64_t r3 = quad_sadj_r3(k, l);
  double* store = Q.work;  // [sub-step][stage][k l]: the stage inputs
  SAdjBufs w;
  if (kLds) {
    w.B0 = Gbb + l * l;
    w.B1 = w.B0 + r3;
    w.B2 = w.B1 + r3;
    double* tl = w.B2 + r3;
    if (Q.T)
      for (int e = t; e < l * l * l; e += nt) tl[e] = Q.T[e];
    w.T = tl;
  } else {
    w.B0 = store + Q.nsub * 4 * kl;
    w.B1 = w.B0 + r3;
    w.B2 = w.B1 + r3;
    w.T = Q.T;
  }
  const double h = Q.h;
  for (int e = t; e < kl; e += nt) {
    const int i = e % k, j = e / k;
    y[e] = Q.S0[i + j * Q.lds0];
    yb[e] = Q.Sb1[i + j * Q.ldsb1];
    double s = 0.0;
    for (int qq = 0; qq < Q.q; ++qq) s += Q.gaT[qq + i * Q.q] * Q.hbT[qq + j * Q.q];
    cb[e] = s;
    cbar[e] = 0.0;
  }
  for (int e = t; e < k * k; e += nt) Gab[e] = 0.0;
  for (int e = t; e < l * l; e += nt) Gbb[e] = 0.0;
  if (Q.W) {
    for (int e = t; e < k * k * k; e += nt) Q.Wb[e] = 0.0;
    for (int e = t; e < l * l * l; e += nt) Q.Tb[e] = 0.0;
  }
  __syncthreads();
  // forward trajectory, keeping every stage input
  for (int64_t st = 0; st < Q.nsub; ++st) {
    double* X4 = store + st * 4 * kl;
    for (int e = t; e < kl; e += nt) {
      X4[e] = y[e];
      x[e] = y[e];
    }
    __syncthreads();
    for (int s = 0; s < 4; ++s) {
      sadj_f(Q, x, fv, cb, tmp, w);
      for (int e = t; e < kl; e += nt) {
        if (s < 3) {
          acc[e] = (s == 0 ? 0.0 : acc[e]) + stage_w(s) * fv[e];
          x[e] = y[e] + stage_c(s) * h * fv[e];
          X4[(s + 1) * kl + e] = x[e];
        } else {
          y[e] = y[e] + (h / 6.0) * (acc[e] + fv[e]);
        }
      }
      __syncthreads();
    }
  }
  // reverse
  for (int64_t st = Q.nsub - 1; st >= 0; --st) {
    const double* X4 = store + st * 4 * kl;
    for (int s = 3; s >= 0; --s) {
      for (int e = t; e < kl; e += nt) {
        x[e] = X4[s * kl + e];
        kb[e] = (h * stage_w(s) / 6.0) * yb[e] + (s < 3 ? stage_c(s) * h * xb[e] : 0.0);
      }
      __syncthreads();
      sadj_rev(Q, x, kb, pb, xb, tmp, Gab, Gbb, cbar, w);
      for (int e = t; e < kl; e += nt) acc[e] = (s == 3 ? yb[e] : acc[e]) + xb[e];
      __syncthreads();
    }
    for (int e = t; e < kl; e += nt) yb[e] = acc[e];
    __syncthreads();
  }
  for (int e = t; e < kl; e += nt) Q.Sb0[e] = yb[e];
  for (int e = t; e < k * k; e += nt) Q.Gab[e] = Gab[e];
  for (int e = t; e < l * l; e += nt) Q.Gbb[e] = Gbb[e];
  for (int e = t; e < Q.q * k; e += nt) {  // gaT_qi: cb_ij = sum_q gaT_qi hbT_qj
    const int qq = e % Q.q, i = e / Q.q;
    double v = 0.0;
    for (int j = 0; j < l; ++j) v += Q.hbT[qq + j * Q.q] * cbar[i + j * k];
    Q.gaTb[e] = v;
  }
  for (int e = t; e < Q.q * l; e += nt) {
    const int qq = e % Q.q, j = e / Q.q;
    double v = 0.0;
    for (int i = 0; i < k; ++i) v += Q.gaT[qq + i * Q.q] * cbar[i + j * k];
    Q.hbTb[e] = v;
  }
}

hipError_t launch_quad_sadj(const QuadSAdjArgs& Q, hipStream_t st) {
  if (Q.k < 1 || Q.l < 1 || Q.k > kQuadMaxRank || Q.l > kQuadMaxRank || Q.q < 0 || Q.q > kMaxForce || !Q.work ||
      Q.nsub < 1 || !Q.S0 || !Q.Sb1 || !Q.Sb0 || !Q.Ga || !Q.Gb || !Q.Gab || !Q.Gbb ||
      (Q.q > 0 && (!Q.gaT || !Q.hbT || !Q.gaTb || !Q.hbTb)) || (Q.W && (!Q.T || !Q.Wb || !Q.Tb)))
    return hipErrorInvalidValue;
  const int small = quad_sadj_small_doubles(Q.k, Q.l);
  const int full = small + 3 * quad_sadj_r3(Q.k, Q.l) + Q.l * Q.l * Q.l;
  if (full <= kSAdjLdsMax)
    hipLaunchKernelGGL(k_quad_sadj<true>, dim3(1), dim3(1024), sizeof(double) * full, st, Q);
  else
    hipLaunchKernelGGL(k_quad_sadj<false>, dim3(1), dim3(1024), sizeof(double) * small, st, Q);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Basis cotangents from the Galerkin data's: tiles of R = 256 / k rows with one
// halo row each side (A, D and D' are tridiagonal).
// ----------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_basis_vjp(BasisVjpArgs Q) {
  __shared__ double xs[256 + 2 * kQuadMaxRank];
  __shared__ double xsig[256 + 2 * kQuadMaxRank];  // X (Gb + Gb')
  __shared__ double dxb[256 + 2 * kQuadMaxRank];   // dxbar_ia = sum_bc Tb_abc X_ib X_ic
  __shared__ double dxs[256];                      // (D X)
  __shared__ double sig[kQuadMaxRank * kQuadMaxRank];
  const int k = Q.k, t = threadIdx.x;
  const int R = 256 / k;
  const int64_t n = Q.rows;
  for (int e = t; e < k * k; e += blockDim.x) {
    const int l = e % k, c = e / k;
    sig[e] = (Q.Gb ? Q.Gb[l + c * k] + Q.Gb[c + l * k] : 0.0) + (Q.Gb2 ? Q.Gb2[l + c * k] + Q.Gb2[c + l * k] : 0.0);
  }
  // the tensor cotangent as the sum of its sources
  auto tb = [&](int64_t idx) { return Q.Tb[idx] + (Q.Tb2 ? Q.Tb2[idx] : 0.0); };
  const int64_t ntiles = (n + R - 1) / R;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t r0 = tile * R;
    __syncthreads();
    stage_rows(xs, Q.X, Q.ldx, r0, R, k, n);
    __syncthreads();
    for (int e = t; e < (R + 2) * k; e += blockDim.x) {
      const int lr = e / k, a = e % k;
      const double* xr = xs + lr * k;
      double v = 0.0, d = 0.0;
      if (Q.Gb || Q.Gb2)
        for (int l = 0; l < k; ++l) v += xr[l] * sig[l + a * k];
      if (Q.Tb)
        for (int b = 0; b < k; ++b) {
          const int64_t ab = (int64_t)k * (a + k * b);
          double in = 0.0;
          for (int c = 0; c < k; ++c) in += tb(ab + c) * xr[c];
          d += xr[b] * in;
        }
      xsig[e] = v;
      dxb[e] = d;
    }
    if (Q.Tb)
      for (int e = t; e < R * k; e += blockDim.x) {
        const int lr = e / k, a = e % k;
        const int64_t i = r0 + lr;
        dxs[e] = i < n ? apply_d(xs, lr, a, k, i, n, Q.dl, Q.dd, Q.du, Q.dper) : 0.0;
      }
    __syncthreads();
    const int lr = t % R, c = t / R;
    const int64_t i = r0 + lr;
    if (c < k && i < n) {
      double v = Q.base ? Q.base[i + c * Q.ldb] : 0.0;
      if (Q.Tb) {
        const double* dr = dxs + lr * k;
        const double* xr = xs + (lr + 1) * k;
        double tv = 0.0;
        for (int a = 0; a < k; ++a) {
          const int64_t ac = (int64_t)k * (a + k * c);  // Tb_{a c x} at ac + x
          double in1 = 0.0, in2 = 0.0;
          for (int x = 0; x < k; ++x) {
            in1 += tb(ac + x) * xr[x];
            in2 += tb(c + (int64_t)k * (a + k * x)) * xr[x];  // Tb_{a x c}
          }
          tv += dr[a] * (in1 + in2);
        }
        v += tv + apply_dt(dxb, lr, c, k, i, n, Q.dl, Q.dd, Q.du, Q.dper);
      }
      if ((Q.Gb || Q.Gb2) && Q.st_d) v += apply_a(xsig, lr, c, k, i, n, Q.st_d, Q.st_e, Q.periodic);
      for (int qq = 0; qq < Q.q; ++qq)
        v += Q.F[i + qq * Q.ldf] * (Q.fb[qq + c * Q.q] + (Q.fb2 ? Q.fb2[qq + c * Q.q] : 0.0));
      if (Q.X2)
        for (int l = 0; l < k; ++l)
          v += Q.X2[i + l * Q.ldx2] * (Q.transY2 ? Q.Y2[c + l * Q.ldy2] : Q.Y2[l + c * Q.ldy2]);
      if (Q.X3)
        for (int l = 0; l < k; ++l)
          v += Q.X3[i + l * Q.ldx3] * (Q.transY3 ? Q.Y3[c + l * Q.ldy3] : Q.Y3[l + c * Q.ldy3]);
      Q.out[i + c * Q.ldo] = v;
    }
  }
}

hipError_t launch_basis_vjp(const BasisVjpArgs& Q, hipStream_t st) {
  if (Q.k < 1 || Q.k > kQuadMaxRank || Q.rows < 1 || !Q.X || !Q.out || Q.q < 0 || Q.q > kMaxForce ||
      (Q.q > 0 && (!Q.F || !Q.fb)) || (Q.Gb && Q.st_d && !Q.st_e) || (Q.X2 && !Q.Y2) || (Q.X3 && !Q.Y3) ||
      (Q.Tb2 && !Q.Tb))
    return hipErrorInvalidValue;
  const int R = 256 / Q.k;
  const int64_t tiles = (Q.rows + R - 1) / R;
  const int grid = (int)std::min<int64_t>(tiles, 1024);
  hipLaunchKernelGGL(k_basis_vjp, dim3(grid), dim3(256), 0, st, Q);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_quad_adj_scales(const double* sumK, const double* sumL, int len,
                                                         const double* Gab0, const double* Ga0, const double* Gab1,
                                                         const double* Ga1, const double* Gbb0, const double* Gb0,
                                                         const double* Gbb1, const double* Gb1, int k, int k1,
                                                         double* out) {
  __shared__ double ra[256], rb[256];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int e = t; e < k * k; e += blockDim.x) {
    a += Gab0[e] * Ga0[e];
    b += Gbb0[e] * Gb0[e];
  }
  for (int e = t; e < k1 * k1; e += blockDim.x) {
    a += Gab1[e] * Ga1[e];
    b += Gbb1[e] * Gb1[e];
  }
  ra[t] = a;
  rb[t] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      ra[t] += ra[t + w];
      rb[t] += rb[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = sumK[len - 1] + ra[0];
    out[1] = sumL[len - 1] + rb[0];
  }
}

hipError_t launch_quad_adj_scales(const double* sumK, const double* sumL, int len, const double* Gab0,
                                  const double* Ga0, const double* Gab1, const double* Ga1, const double* Gbb0,
                                  const double* Gb0, const double* Gbb1, const double* Gb1, int k, int k1,
                                  double* out, hipStream_t st) {
  if (k < 1 || k > kQuadMaxRank || k1 < 1 || k1 > kQuadMaxRank || len < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_quad_adj_scales, dim3(1), dim3(256), 0, st, sumK, sumL, len, Gab0, Ga0, Gab1, Ga1, Gbb0, Gb0,
                     Gbb1, Gb1, k, k1, out);
  return hipGetLastError();
}

}  // namespace lri
