// Matrix-ODE sub-steps for right-hand sides with a quadratic term,
//
//     F(Y) = A Y + Y B' + G H' - (D Y) o Y          (o: elementwise product)
//
// A, B symmetric tridiagonal and D a general tridiagonal acting on the rows
// (e.g. viscous Burgers: A = nu Laplacian, D = central gradient, both
// periodic; each column one parameter sample).  With U, V orthonormal and
// frozen during a sub-step, the reference's default right-hand sides
// (projector_splitting.jl:52-80, unconventional.jl:51-79,
// rank_adaptive_unconventional.jl:59-86) are
//
//     K' = A K + K (V'BV) + G (H'V) - N,  N_ic = sum_ab (D K)_ia K_ib T_abc,
//     L' = B L + L (U'AU) + H (G'U) - Q,  Q_jc = sum_ab L_ja L_jb W_abc,
//     S' = sgn ((U'AU) S + S (V'BV) + (U'G)(H'V)' - P),
//          P_cd = sum_{e f a b} W_efc S_ea S_fb T_abd,
//
// with the third-order tensors T_abc = sum_j V_ja V_jb V_jc and
// W_abc = sum_i (D U)_ia U_ib U_ic.  The sub-problems are no longer linear, so
// every one of the N classical RK4 steps evaluates them: the K and L steps one
// launch per RK4 stage over all rows (a stage needs the neighbouring rows of
// the previous one), the S step in one workgroup.
#include <algorithm>

#include "lri_common.h"
#include "lri_kernels.h"

namespace lri {

namespace {

// RK4 stage weights: y_{s+1} input = y0 + c_s h k_s, accumulation weight w_s
__device__ __forceinline__ double stage_c(int s) { return s == 2 ? 1.0 : 0.5; }
__device__ __forceinline__ double stage_w(int s) { return (s == 1 || s == 2) ? 2.0 : 1.0; }

// row i of X (n x k, ld n), with periodic wrap or zero outside
__device__ __forceinline__ int64_t wrap_row(int64_t i, int64_t n, int periodic, bool* ok) {
  if (i < 0) {
    *ok = periodic != 0;
    return i + n;
  }
  if (i >= n) {
    *ok = periodic != 0;
    return i - n;
  }
  *ok = true;
  return i;
}

}  // namespace

// ----------------------------------------------------------------------------
// Partial third-order tensors: workgroup g sums the rows of chunk g,
//   part[g][c + k (a + k b)] = sum_{r in chunk} X_ra Y_rb Y_rc,
// X = D Y (the tridiagonal D applied to Y) when Q.dd != nullptr, else X = Y.
// ----------------------------------------------------------------------------

constexpr int kTensorRows = 64;

__global__ __launch_bounds__(256) void k_tensor3_part(TensorArgs Q) {
  __shared__ double xs[kTensorRows * kQuadMaxRank];
  __shared__ double ys[kTensorRows * kQuadMaxRank];
  const int k = Q.k, t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * kTensorRows;
  for (int e = t; e < kTensorRows * k; e += blockDim.x) {
    const int lr = e % kTensorRows, c = e / kTensorRows;
    const int64_t i = r0 + lr;
    double y = 0.0, x = 0.0;
    if (i < Q.rows) {
      y = Q.Y[i + c * Q.ldy];
      if (Q.dd) {
        bool okl, oku;
        const int64_t il = wrap_row(i - 1, Q.rows, Q.dper, &okl);
        const int64_t iu = wrap_row(i + 1, Q.rows, Q.dper, &oku);
        x = Q.dd[i] * y;
        if (okl) x += Q.dl[i] * Q.Y[il + c * Q.ldy];
        if (oku) x += Q.du[i] * Q.Y[iu + c * Q.ldy];
      } else {
        x = y;
      }
    }
    xs[lr * k + c] = x;
    ys[lr * k + c] = y;
  }
  __syncthreads();
  const int k3 = k * k * k;
  double* out = Q.part + (int64_t)blockIdx.x * k3;
  for (int e = t; e < k3; e += blockDim.x) {
    const int c = e % k, a = (e / k) % k, b = e / (k * k);
    double s = 0.0;
    for (int lr = 0; lr < kTensorRows; ++lr) s += xs[lr * k + a] * ys[lr * k + b] * ys[lr * k + c];
    out[e] = s;
  }
}

// out[e] = sum over chunks, in chunk order
__global__ __launch_bounds__(256) void k_tensor3_sum(const double* __restrict__ part, int nchunks, int k3,
                                                     double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k3) return;
  double s = 0.0;
  for (int g = 0; g < nchunks; ++g) s += part[(int64_t)g * k3 + e];
  out[e] = s;
}

hipError_t launch_tensor3(const TensorArgs& Q, double* out, hipStream_t st) {
  if (Q.k < 1 || Q.k > kQuadMaxRank || Q.rows < 1 || !Q.part) return hipErrorInvalidValue;
  const int nchunks = (int)((Q.rows + kTensorRows - 1) / kTensorRows);
  hipLaunchKernelGGL(k_tensor3_part, dim3(nchunks), dim3(256), 0, st, Q);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int k3 = Q.k * Q.k * Q.k;
  hipLaunchKernelGGL(k_tensor3_sum, dim3((k3 + 255) / 256), dim3(256), 0, st, (const double*)Q.part, nchunks, k3,
                     out);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// One RK4 stage of the K or L sub-problem over all rows.  Workgroup: a tile of
// 16 rows (4 waves).  The tile's rows (with one halo row on each side) are
// staged in LDS, then
//   f_ic = (A Xs)_ic + (Xs M)_ic + (g hv)_ic - N_ic,  N_ic = sum_ab (P Xs)_ia Xs_ib T_abc
// (P = D when Q.dd, else the identity), and the RK4 bookkeeping of stage s:
//   s < 3: Acc (=, +=) w_s f, Xn = X0 + c_s h f;   s = 3: X0 += h/6 (Acc + f).
// The quadratic term is a product on the FP64 matrix cores,
//   N (16 x k) = Z (16 x k^2) T (k^2 x k),  Z_i,(a + k b) = (P Xs)_ia Xs_ib,
// T read once per tile in the k^2 order of its storage T[c + k (a + k b)]:
// wave w takes column block w % nct (nct = ceil(k / 16)) and the w / nct-th of
// the 4 / nct slices of the k^2 reduction; the slices are summed through LDS.
// ----------------------------------------------------------------------------

constexpr int kStageRows = 16;

__global__ __launch_bounds__(256) void k_quad_stage(QuadStageArgs Q) {
  __shared__ double xs[(kStageRows + 2) * kQuadMaxRank];  // rows r0 - 1 .. r0 + 16: xs[lr * k + c]
  __shared__ double ps[kStageRows * kQuadMaxRank];        // (P Xs) rows r0 .. r0 + 15
  __shared__ double nl[4][kStageRows * kQuadMaxRank];     // slices of N, [row * 32 + c]
  const int k = Q.k, t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, lr = lane & 15, lk = lane >> 4;
  const int64_t r0 = (int64_t)blockIdx.x * kStageRows;
  const int64_t n = Q.n;
  // a row outside [0, n) takes its periodic image (used only by a periodic stencil)
  for (int e = t; e < (kStageRows + 2) * k; e += blockDim.x) {
    const int lrow = e % (kStageRows + 2), c = e / (kStageRows + 2);
    int64_t ir = r0 + lrow - 1;
    if (ir < 0) ir += n;
    else if (ir >= n) ir -= n;
    xs[lrow * k + c] = (ir >= 0 && ir < n) ? Q.Xs[ir + c * Q.ld] : 0.0;
  }
  __syncthreads();
  for (int e = t; e < kStageRows * k; e += blockDim.x) {
    const int lrow = e % kStageRows, c = e / kStageRows;
    const int64_t i = r0 + lrow;
    double p = 0.0;
    if (i < n) {
      const double xm = xs[lrow * k + c], x0 = xs[(lrow + 1) * k + c], xp = xs[(lrow + 2) * k + c];
      const bool hasl = i > 0 || Q.dper, hasu = i < n - 1 || Q.dper;
      p = Q.dd ? (Q.dd[i] * x0 + (hasl ? Q.dl[i] * xm : 0.0) + (hasu ? Q.du[i] * xp : 0.0)) : x0;
    }
    ps[lrow * k + c] = p;
  }
  __syncthreads();
  const int nct = (k + 15) / 16, nch = 4 / nct;
  if (Q.T) {
    const int ct = wave % nct, ch = wave / nct;
    const int kk2 = k * k;
    const int clen = (kk2 + 4 * nch - 1) / (4 * nch) * 4;
    const int kbeg = ch * clen, kend = min(kbeg + clen, kk2);
    const int c = ct * 16 + lr;
    f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
    // the tensor operand of the next k-step is loaded before this one's product
    double bn = (kbeg + lk < kend && c < k) ? Q.T[c + (int64_t)k * (kbeg + lk)] : 0.0;
    for (int kk0 = kbeg; kk0 < kend; kk0 += 4) {
      const int kk = kk0 + lk;
      const double b = bn;
      bn = (kk + 4 < kend && c < k) ? Q.T[c + (int64_t)k * (kk + 4)] : 0.0;
      const double a = kk < kend ? ps[lr * k + kk % k] * xs[(lr + 1) * k + kk / k] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) nl[ch][(4 * u + lk) * kQuadMaxRank + c] = acc[u];
  }
  __syncthreads();
  for (int e = t; e < kStageRows * k; e += blockDim.x) {
    const int lrow = e % kStageRows, c = e / kStageRows;
    const int64_t i = r0 + lrow;
    if (i >= n) continue;
    const double* xrow = xs + (lrow + 1) * k;
    const double xm = xs[lrow * k + c], x0 = xrow[c], xp = xs[(lrow + 2) * k + c];
    double f = 0.0;
    // A Xs: symmetric tridiagonal, e[i] couples rows i and i + 1 (e[n-1]: n-1 and 0 when periodic)
    if (Q.st_d) {
      f = Q.st_d[i] * x0;
      if (i > 0) f += Q.st_e[i - 1] * xm;
      else if (Q.periodic) f += Q.st_e[n - 1] * xm;
      if (i < n - 1) f += Q.st_e[i] * xp;
      else if (Q.periodic) f += Q.st_e[n - 1] * xp;
    }
    if (Q.M)
      for (int l = 0; l < k; ++l) f += xrow[l] * Q.M[l + c * k];
    for (int qq = 0; qq < Q.q; ++qq) f += Q.g[i + qq * Q.ldg] * Q.hv[qq + c * Q.q];
    if (Q.T)
      for (int ch = 0; ch < nch; ++ch) f -= nl[ch][lrow * kQuadMaxRank + c];
    const int64_t o = i + c * Q.ld;
    const int s = Q.stage;
    if (s < 3) {
      Q.Acc[o] = (s == 0 ? 0.0 : Q.Acc[o]) + stage_w(s) * f;
      Q.Xn[o] = Q.X0[o] + stage_c(s) * Q.h * f;
    } else {
      Q.X0[o] = Q.X0[o] + (Q.h / 6.0) * (Q.Acc[o] + f);
    }
  }
}

hipError_t launch_quad_stage(const QuadStageArgs& Q, hipStream_t st) {
  if (Q.k < 1 || Q.k > kQuadMaxRank || Q.n < 1 || Q.stage < 0 || Q.stage > 3 || Q.q < 0 || Q.q > kMaxForce)
    return hipErrorInvalidValue;
  const int64_t blocks = (Q.n + kStageRows - 1) / kStageRows;
  hipLaunchKernelGGL(k_quad_stage, dim3((unsigned)blocks), dim3(256), 0, st, Q);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// S sub-step: N RK4 steps of
//   S' = sgn (Ga S + S Gb + cb - P(S)),  P_cd = sum_{a b} Z_abc T_abd,
//   Z_abc = sum_e S_ea Y_ebc,  Y_ebc = sum_f S_fb W_efc,
// S k x l in LDS, Y (k x l x k) and Z (l x l x k) in global scratch; one workgroup.
// ----------------------------------------------------------------------------

// kLds: the state, its RK4 work, the tensors W, T and the contractions Y, Z all
// in LDS (quad_s_lds_doubles(k, l) <= kQuadSLds), else Y, Z in Q.work and W, T read
// from global memory.
constexpr int kQuadSLds = 20400;  // 159 KB of the 160 KB LDS

__host__ __device__ inline int quad_s_lds_doubles(int k, int l) {
  return 5 * k * l + k * l * k + l * l * k + k * k * k + l * l * l;
}

template <bool kLds>
__global__ __launch_bounds__(1024) void k_quad_sstep(QuadSArgs Q) {
  __shared__ double st_small[kLds ? 1 : 5 * kQuadMaxRank * kQuadMaxRank];
  extern __shared__ double dyn[];
  const int k = Q.k, l = Q.l, t = threadIdx.x, nt = blockDim.x;
  const int kl = k * l;
  double* base = kLds ? dyn : st_small;
  double* y0 = base;
  double* ys = y0 + kl;
  double* acc = ys + kl;
  double* cb = acc + kl;
  double* pv = cb + kl;                          // the quadratic term P of a stage
  double* Y = kLds ? pv + kl : Q.work;           // k x l x k: Y[c + k (e + k b)]
  double* Z = Y + k * l * k;                     // l x l x k: Z[c + k (a + l b)]
  const double* W = Q.W;
  const double* T = Q.T;
  if (kLds && Q.W) {
    double* w = Z + l * l * k;
    double* tt = w + k * k * k;
    for (int e = t; e < k * k * k; e += nt) w[e] = Q.W[e];
    for (int e = t; e < l * l * l; e += nt) tt[e] = Q.T[e];
    W = w;
    T = tt;
  }
  for (int e = t; e < kl; e += nt) {
    const int i = e % k, j = e / k;
    y0[e] = Q.transS0 ? Q.S0[j + i * Q.lds0] : Q.S0[i + j * Q.lds0];
    double s = 0.0;
    for (int qq = 0; qq < Q.q; ++qq) s += Q.gaT[qq + i * Q.q] * Q.hbT[qq + j * Q.q];
    cb[e] = s;
  }
  __syncthreads();
  const double h = Q.h;
  for (int64_t step = 0; step < Q.nsub; ++step) {
    for (int e = t; e < kl; e += nt) ys[e] = y0[e];
    __syncthreads();
    for (int s = 0; s < 4; ++s) {
      if (Q.W) {
        // on the matrix cores: Y_(c + k e), b = sum_f W_(c + k e), f S_fb;  Z_(c, a), b = sum_e Y_(c, e), b S_ea;
        // P_ij = sum_(a + l b) Z_i, (a + l b) T_(a + l b), j  (into pv)
        const int kk2 = k * k;
        wave_gemm(kk2, l, k, 1, [&](int, int m, int f) { return W[m + kk2 * f]; },
                  [&](int, int f, int b) { return ys[f + b * k]; },
                  [&](int, int m, int b, double v) { Y[m + kk2 * b] = v; });
        __syncthreads();
        wave_gemm(k, l, k, l, [&](int b, int c, int e) { return Y[c + k * e + kk2 * b]; },
                  [&](int, int e, int a) { return ys[e + a * k]; },
                  [&](int b, int c, int a, double v) { Z[c + k * a + k * l * b] = v; });
        __syncthreads();
        wave_gemm(k, l, l * l, 1, [&](int, int i, int ab) { return Z[i + k * ab]; },
                  [&](int, int ab, int j) { return T[j + l * ab]; },
                  [&](int, int i, int j, double v) { pv[i + j * k] = v; });
        __syncthreads();
      }
      double fv[(kQuadMaxRank * kQuadMaxRank + 1023) / 1024];
      int cnt = 0;
      for (int e = t; e < kl; e += nt, ++cnt) {
        const int i = e % k, j = e / k;
        double v = cb[e];
        for (int m = 0; m < k; ++m) v += Q.Ga[i + m * k] * ys[m + j * k];
        for (int m = 0; m < l; ++m) v += ys[i + m * k] * Q.Gb[m + j * l];
        if (Q.W) v -= pv[e];
        fv[cnt] = Q.sgn * v;
      }
      __syncthreads();  // every stage input read before it is replaced
      cnt = 0;
      for (int e = t; e < kl; e += nt, ++cnt) {
        const double f = fv[cnt];
        if (s < 3) {
          acc[e] = (s == 0 ? 0.0 : acc[e]) + stage_w(s) * f;
          ys[e] = y0[e] + stage_c(s) * h * f;
        } else {
          y0[e] = y0[e] + (h / 6.0) * (acc[e] + f);
        }
      }
      __syncthreads();
    }
  }
  for (int e = t; e < kl; e += nt) Q.S1[(e % k) + (e / k) * Q.lds1] = y0[e];
}

hipError_t launch_quad_sstep(const QuadSArgs& Q, hipStream_t st) {
  if (Q.k < 1 || Q.l < 1 || Q.k > kQuadMaxRank || Q.l > kQuadMaxRank || Q.q < 0 || Q.q > kMaxForce || !Q.work)
    return hipErrorInvalidValue;
  const int need = quad_s_lds_doubles(Q.k, Q.l);
  if (need <= kQuadSLds)
    hipLaunchKernelGGL(k_quad_sstep<true>, dim3(1), dim3(1024), sizeof(double) * need, st, Q);
  else
    hipLaunchKernelGGL(k_quad_sstep<false>, dim3(1), dim3(1024), 0, st, Q);
  return hipGetLastError();
}

}  // namespace lri
