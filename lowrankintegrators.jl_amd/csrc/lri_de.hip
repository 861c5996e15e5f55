// Matrix-ODE (MatrixDEProblem) sub-steps for the operator
//
//     F(Y) = A Y + Y B' + sum_k g_k h_k'
//
// with A (n x n) and B (m x m) symmetric tridiagonal (optionally periodic) and a
// rank-q forcing.  With U, V fixed during a sub-step the reference's default
// right-hand sides (projector_splitting.jl:52-80, unconventional.jl:51-79,
// rank_adaptive_unconventional.jl:59-86) become linear:
//
//     K' = A K + K Bt + g (V'h)',   Bt = V'BV          (K step)
//     L' = B L + L At + h (U'g)',   At = U'AU          (L step)
//     S' = sgn (At S + S Bt + (U'g)(V'h)')             (S step; sgn = -1 in KSL)
//
// Bt is symmetric, so with Bt = W diag(beta) W' the columns of Kbar = K W
// decouple: kbar_j' = (A + beta_j) kbar_j + g (h'V W)_j.  Classical RK4 commutes
// with a constant orthogonal change of variables, so N RK4 steps on Kbar
// followed by K = Kbar W' equal N RK4 steps on K, which is what a fixed-step
// RK4 sub-integrator does in the reference's K/L/S sub-problems.  Each column
// then runs all of its RK4 stages inside one workgroup.  The S step is
// diagonal in the product eigenbasis and its N RK4 steps are applied in closed
// form entry by entry.
#include <algorithm>

#include "lri_common.h"
#include "lri_kernels.h"

namespace lri {

// ----------------------------------------------------------------------------
// RK4 on independent columns.  Workgroup j owns column j; thread t owns rows
// [t*RPT, t*RPT + RPT) in registers and only the first and last of them go
// through LDS (double-buffered by stage, so one barrier per stage).  A stage
// is evaluated in ascending row order in place: row u needs the old values of
// rows u-1 (kept in `prev`) and u+1 (not yet overwritten).
// REG keeps the row coefficients in registers; otherwise they are re-read from
// the (cache-resident) operator arrays at each stage.
// ----------------------------------------------------------------------------

template <int RPT, bool REG>
__global__ __launch_bounds__(1024) void k_rk4_columns(RK4ColArgs P0, RK4ColArgs P1) {
  __shared__ double s_first[2][1024], s_last[2][1024];
  // workgroups [0, P0.ncols) integrate the columns of P0, the rest those of P1
  const bool second = (int)blockIdx.x >= P0.ncols;
  const RK4ColArgs P = second ? P1 : P0;
  const int t = threadIdx.x;
  const int j = second ? (int)blockIdx.x - P0.ncols : (int)blockIdx.x;
  const int64_t n = P.n;
  const int64_t r0 = (int64_t)t * RPT;
  const int last_t = (int)((n - 1) / RPT);
  const bool active = t <= last_t;
  const int nlast = (int)(n - (int64_t)last_t * RPT);
  const int nmine = !active ? 0 : (t == last_t ? nlast : RPT);
  const int tn = (t == last_t) ? 0 : t + 1;  // holder of the row after my last row
  const int tp = (t == 0) ? last_t : t - 1;  // holder of the row before my first row
  const bool has_dn = active && (t < last_t || P.periodic);
  const bool has_up = active && (t > 0 || P.periodic);
  const double mu = P.mu[j];
  const double h = P.h, h2 = 0.5 * h, h6 = h / 6.0;
  double cf[kMaxForce];
#pragma unroll
  for (int k = 0; k < kMaxForce; ++k) cf[k] = k < P.q ? P.coef[k + j * P.q] : 0.0;
  // e of the row above my first row
  double eup = 0.0;
  if (has_up) eup = r0 > 0 ? P.st_e[r0 - 1] : P.st_e[n - 1];

  // row pointers of this thread (dereferenced only for its own rows)
  const double* pd = P.st_d + (active ? r0 : 0);
  const double* pe = P.st_e + (active ? r0 : 0);
  const double* pg = P.g ? P.g + (active ? r0 : 0) : nullptr;
  // e of the last row: zero unless periodic
  const bool e_last = P.periodic != 0;
  auto load_coef = [&](int u, double& dd, double& ee, double& ff) {
    if (u < nmine) {
      dd = pd[u] + mu;
      ee = (r0 + u + 1 < n || e_last) ? pe[u] : 0.0;
      double s = 0.0;
      for (int k = 0; k < P.q; ++k) s += pg[u + k * P.ldg] * cf[k];
      ff = s;
    } else {
      dd = ee = ff = 0.0;
    }
  };

  double d[REG ? RPT : 1], e[REG ? RPT : 1], f[REG ? RPT : 1];
  if (REG) {
#pragma unroll
    for (int u = 0; u < RPT; ++u) load_coef(u, d[REG ? u : 0], e[REG ? u : 0], f[REG ? u : 0]);
  }

  double y[RPT], ks[RPT], x[RPT];
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    y[u] = u < nmine ? P.X[r0 + u + (int64_t)j * P.ldx] : 0.0;
    x[u] = y[u];
    ks[u] = 0.0;
  }
  int buf = 0;
  for (int64_t s = 0; s < P.nsub; ++s) {
#pragma unroll 1
    for (int stage = 0; stage < 4; ++stage) {
      // stage input x = y + c_stage * k, accumulated weight w_stage
      const double cs = stage < 2 ? h2 : h;
      const double ws = (stage == 0) ? 1.0 : 2.0;
      s_first[buf][t] = x[0];
      double xl = x[RPT - 1];
#pragma unroll
      for (int u = 0; u < RPT; ++u)
        if (u == nmine - 1) xl = x[u];
      s_last[buf][t] = xl;
      __syncthreads();
      const double above = has_up ? s_last[buf][tp] : 0.0;
      const double below = has_dn ? s_first[buf][tn] : 0.0;
      buf ^= 1;
      double prev = above, e_prev = eup;
#pragma unroll
      for (int u = 0; u < RPT; ++u) {
        double dd, ee, ff;
        if (REG) {
          dd = d[REG ? u : 0];
          ee = e[REG ? u : 0];
          ff = f[REG ? u : 0];
        } else {
          load_coef(u, dd, ee, ff);
        }
        const bool real = u < nmine;
        const double xu = x[u];
        const double xd = (u + 1 < nmine) ? x[u + 1 < RPT ? u + 1 : RPT - 1] : below;
        const double k = real ? dd * xu + e_prev * prev + ee * xd + ff : 0.0;
        prev = xu;
        e_prev = ee;
        if (stage < 3) {
          ks[u] = stage == 0 ? k : ks[u] + ws * k;
          x[u] = y[u] + cs * k;
        } else {
          y[u] = y[u] + h6 * (ks[u] + k);
          x[u] = y[u];
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < RPT; ++u)
    if (u < nmine) P.Y[r0 + u + (int64_t)j * P.ldy] = y[u];
}

hipError_t launch_rk4_columns(const RK4ColArgs& P0, const RK4ColArgs& P1, hipStream_t st) {
  if (current_recorder()) return hipErrorNotSupported;  // no batched form
  const RK4ColArgs* ps[2] = {&P0, &P1};
  int64_t nmax = 0;
  for (const RK4ColArgs* P : ps) {
    if (P->ncols < 1) continue;
    if (P->nsub < 1 || P->n < 1 || P->n > 16384 || P->q < 0 || P->q > kMaxForce) return hipErrorInvalidValue;
    nmax = P->n > nmax ? P->n : nmax;
  }
  if (P0.ncols < 0 || P1.ncols < 0) return hipErrorInvalidValue;
  const int ncols = P0.ncols + P1.ncols;
  if (ncols < 1) return hipSuccess;
  int rpt = 4;
  while ((int64_t)rpt * 1024 < nmax) rpt *= 2;
  int nt = (int)((nmax + rpt - 1) / rpt);
  nt = (nt + 63) & ~63;
  const dim3 grid((unsigned)ncols), blk((unsigned)nt);
  if (rpt == 4)
    hipLaunchKernelGGL((k_rk4_columns<4, true>), grid, blk, 0, st, P0, P1);
  else if (rpt == 8)
    hipLaunchKernelGGL((k_rk4_columns<8, true>), grid, blk, 0, st, P0, P1);
  else
    hipLaunchKernelGGL((k_rk4_columns<16, false>), grid, blk, 0, st, P0, P1);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// CW = op(C) W (k x k) and coef = FX W (q x k), one workgroup
// ----------------------------------------------------------------------------

// CW = op(C) W (k x k) and coef = FX W (q x k); kLds stages op(C) and W in LDS
// (k <= 64), where the products read them at an odd stride.
template <bool kLds>
__device__ __forceinline__ void de_prep_body(const double* __restrict__ C, int ldc, int transC,
                                             const double* __restrict__ W, const double* __restrict__ FX, int k,
                                             int q, double* __restrict__ CW, double* __restrict__ coef) {
  extern __shared__ double sm[];
  const int t = threadIdx.x, nt = blockDim.x;
  const int ld = k | 1;
  if (kLds) {
    double* Cs = sm;           // op(C), k x k, ld
    double* Ws = sm + ld * k;  // W, k x k, ld
    for (int e = t; e < k * k; e += nt) {
      const int i = e % k, c = e / k;
      Cs[i + c * ld] = transC ? C[c + i * ldc] : C[i + c * ldc];
      Ws[i + c * ld] = W[e];
    }
    __syncthreads();
    for (int e = t; e < k * k; e += nt) {
      const int i = e % k, jj = e / k;
      double s = 0.0;
      for (int c = 0; c < k; ++c) s += Cs[i + c * ld] * Ws[c + jj * ld];
      CW[e] = s;
    }
    for (int e = t; e < q * k; e += nt) {
      const int qq = e % q, jj = e / q;
      double s = 0.0;
      for (int c = 0; c < k; ++c) s += FX[qq + c * q] * Ws[c + jj * ld];
      coef[e] = s;
    }
    return;
  }
  for (int e = t; e < k * k; e += nt) {
    const int i = e % k, jj = e / k;
    double s = 0.0;
    for (int c = 0; c < k; ++c) s += (transC ? C[c + i * ldc] : C[i + c * ldc]) * W[c + jj * k];
    CW[e] = s;
  }
  for (int e = t; e < q * k; e += nt) {
    const int qq = e % q, jj = e / q;
    double s = 0.0;
    for (int c = 0; c < k; ++c) s += FX[qq + c * q] * W[c + jj * k];
    coef[e] = s;
  }
}

struct DePrepLaunch {
  const double* C;
  int ldc, transC;
  const double *W, *FX;
  int k, q;
  double *CW, *coef;
};
template <bool kLds>
__global__ __launch_bounds__(256) void k_de_prep(DePrepLaunch L) {
  de_prep_body<kLds>(L.C, L.ldc, L.transC, L.W, L.FX, L.k, L.q, L.CW, L.coef);
}
template <bool kLds>
__global__ __launch_bounds__(256) void k_de_prep_batch(const unsigned char* __restrict__ tab,
                                                       const int* __restrict__ start, int count) {
  const int i = batch_member(start, count, blockIdx.x);
  const DePrepLaunch& L = *reinterpret_cast<const DePrepLaunch*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes);
  de_prep_body<kLds>(L.C, L.ldc, L.transC, L.W, L.FX, L.k, L.q, L.CW, L.coef);
}

hipError_t launch_de_prep(const double* C, int ldc, int transC, const double* W, const double* FX, int k, int q,
                          double* CW, double* coef, hipStream_t st) {
  if (k < 1 || k > kMaxAug || q < 0 || q > kMaxForce) return hipErrorInvalidValue;
  const DePrepLaunch L{C, ldc, transC, W, FX, k, q, CW, coef};
  const bool lds = k <= 64;
  const size_t bytes = lds ? sizeof(double) * 2 * (size_t)(k | 1) * k : 0;
  if (record_launch(KID_DE_PREP, lds ? 1 : 0, 1, 256, bytes, L)) return hipSuccess;
  if (lds)
    hipLaunchKernelGGL(k_de_prep<true>, dim3(1), dim3(256), bytes, st, L);
  else
    hipLaunchKernelGGL(k_de_prep<false>, dim3(1), dim3(256), 0, st, L);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// S sub-step in the product eigenbasis, one workgroup (k, l <= 128):
//   Sbar = Wa' S0 Wb,  cbar = (Wa' U'g)(Wb' V'h)',
//   s_ij' = sgn ((alpha_i + beta_j) s_ij + cbar_ij)  for N RK4 steps of size h,
//   S1 = Wa Sbar Wb'.
// One RK4 step on y' = a y + b maps y -> R y + h Pz b with z = h a,
// R = 1 + z + z^2/2 + z^3/6 + z^4/24, Pz = 1 + z/2 + z^2/6 + z^3/24; N steps are
// the N-fold composition of this affine map, formed by repeated squaring.
// ----------------------------------------------------------------------------

// N classical RK4 steps of size h on y' = a y + b (z = h a) map y0 to
// Ra y0 + Qb b: one step is y -> R(z) y + h Pz(z) b with the stability
// polynomial R = 1 + z + z^2/2 + z^3/6 + z^4/24 and Pz = 1 + z/2 + z^2/6 + z^3/24,
// and the N-fold composition is formed by repeated squaring.
__device__ __forceinline__ void rk4_power(double z, double h, int64_t N, double& Ra, double& Qb) {
  const double R = 1.0 + z * (1.0 + z * (0.5 + z * (1.0 / 6.0 + z / 24.0)));
  const double Pz = 1.0 + z * (0.5 + z * (1.0 / 6.0 + z / 24.0));
  double Rm = R, Qm = h * Pz;  // one step (b = 1)
  Ra = 1.0;
  Qb = 0.0;
  for (int64_t nn = N; nn > 0; nn >>= 1) {
    if (nn & 1) {
      Qb = Rm * Qb + Qm;
      Ra = Rm * Ra;
    }
    Qm = Rm * Qm + Qm;
    Rm = Rm * Rm;
  }
}

// S <- N RK4 steps of S' = sgn (Ga S + S Gb' + ga hb') in the product eigenbasis
// (Ga = Wa diag(alpha) Wa', Gb = Wb diag(beta) Wb'): Sbar = Wa' op(S0) Wb, each
// entry advanced in closed form, S1 = Wa Sbar Wb'.  kLds: the bases and the
// work matrices sit in LDS (k, l <= 64), else in P.work.
template <bool kLds>
__device__ __forceinline__ void s_substep_body(const SStepArgs& P) {
  extern __shared__ double sm[];
  const int k = P.k, l = P.l, q = P.q, t = threadIdx.x, nt = blockDim.x;
  const double* Wa = P.Wa;
  const double* Wb = P.Wb;
  const int lwa = kLds ? (k | 1) : k;  // odd stride in LDS: Wa is also read along rows
  double* Sb = P.work;                 // k x l
  if (kLds) {
    double* wa = sm;
    double* wb = wa + lwa * k;
    for (int e = t; e < k * k; e += nt) wa[e % k + (e / k) * lwa] = P.Wa[e];
    for (int e = t; e < l * l; e += nt) wb[e] = P.Wb[e];
    Wa = wa;
    Wb = wb;
    Sb = wb + l * l;
  }
  double* T = Sb + k * l;   // k x l
  double* gh = T + k * l;   // k x q: Wa' (G' U)'
  double* hh = gh + k * q;  // l x q: Wb' (H' V)'
  for (int e = t; e < k * l; e += nt) {
    const int i = e % k, jj = e / k;
    Sb[e] = P.transS0 ? P.S0[jj + i * P.lds0] : P.S0[i + jj * P.lds0];
  }
  for (int e = t; e < (k + l) * q; e += nt) {
    if (e < k * q) {
      const int i = e % k, qq = e / k;
      double s = 0.0;
      for (int c = 0; c < k; ++c) s += P.Wa[c + i * k] * P.gaT[qq + c * q];
      gh[e] = s;
    } else {
      const int f = e - k * q, jj = f % l, qq = f / l;
      double s = 0.0;
      for (int c = 0; c < l; ++c) s += P.Wb[c + jj * l] * P.hbT[qq + c * q];
      hh[f] = s;
    }
  }
  __syncthreads();
  // T = op(S0) Wb
  for (int e = t; e < k * l; e += nt) {
    const int i = e % k, jj = e / k;
    double s = 0.0;
    for (int c = 0; c < l; ++c) s += Sb[i + c * k] * Wb[c + jj * l];
    T[e] = s;
  }
  __syncthreads();
  // Sbar = Wa' T, then the RK4 flow per entry
  for (int e = t; e < k * l; e += nt) {
    const int i = e % k, jj = e / k;
    double s = 0.0;
    for (int c = 0; c < k; ++c) s += Wa[c + i * lwa] * T[c + jj * k];
    double cb = 0.0;
    for (int qq = 0; qq < q; ++qq) cb += gh[i + qq * k] * hh[jj + qq * l];
    const double a = P.sgn * (P.alpha[i] + P.beta[jj]);
    double Ra, Qb;
    rk4_power(P.h * a, P.h, P.nsub, Ra, Qb);
    Sb[e] = Ra * s + Qb * (P.sgn * cb);
  }
  __syncthreads();
  // T = Sbar Wb'
  for (int e = t; e < k * l; e += nt) {
    const int i = e % k, jj = e / k;
    double s = 0.0;
    for (int c = 0; c < l; ++c) s += Sb[i + c * k] * Wb[jj + c * l];
    T[e] = s;
  }
  __syncthreads();
  // S1 = Wa T
  for (int e = t; e < k * l; e += nt) {
    const int i = e % k, jj = e / k;
    double s = 0.0;
    for (int c = 0; c < k; ++c) s += Wa[i + c * lwa] * T[c + jj * k];
    P.S1[i + jj * P.lds1] = s;
  }
}

template <bool kLds>
__global__ __launch_bounds__(1024) void k_s_substep(SStepArgs P) {
  s_substep_body<kLds>(P);
}
template <bool kLds>
__global__ __launch_bounds__(1024) void k_s_substep_batch(const unsigned char* __restrict__ tab,
                                                          const int* __restrict__ start, int count) {
  const int i = batch_member(start, count, blockIdx.x);
  s_substep_body<kLds>(*reinterpret_cast<const SStepArgs*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes));
}

hipError_t launch_s_substep(const SStepArgs& P, hipStream_t st) {
  if (P.k < 1 || P.l < 1 || P.k > kMaxAug || P.l > kMaxAug || P.q < 0 || P.q > kMaxForce || !P.work)
    return hipErrorInvalidValue;
  if (P.k <= 64 && P.l <= 64) {
    const size_t lds = sizeof(double) * ((size_t)(P.k | 1) * P.k + (size_t)P.l * P.l + 2 * (size_t)P.k * P.l +
                                         (size_t)(P.k + P.l) * P.q);
    const unsigned th = P.k * P.l > 1024 ? 1024 : 256;
    if (!record_launch(KID_SSTEP, 1, 1, th, lds, P)) hipLaunchKernelGGL(k_s_substep<true>, dim3(1), dim3(th), lds, st, P);
  } else if (!record_launch(KID_SSTEP, 0, 1, 1024, 0, P)) {
    hipLaunchKernelGGL(k_s_substep<false>, dim3(1), dim3(1024), 0, st, P);
  }
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Constant-coefficient operators T = d I + e (S + S'), Dirichlet or periodic,
// have an analytic orthonormal eigenbasis (the discrete sine basis, or the real
// Fourier basis).  In it the column problems of a K/L sub-step decouple
// entrywise, zhat_ij' = (lambda_i + mu_j) zhat_ij + bhat_ij, and their N RK4
// steps are applied in closed form (rk4_power): two n x n x k products replace
// the N * 4 stencil stages of k_rk4_columns with the same result up to rounding.
// ----------------------------------------------------------------------------

// Q (n x n, column-major) and lambda: Dirichlet
//   Q(i, k) = sqrt(2 / (n + 1)) sin(pi (i + 1)(k + 1) / (n + 1)),
//   lambda_k = d + 2 e cos(pi (k + 1) / (n + 1));
// periodic (n >= 3): column 0 = 1 / sqrt(n) (lambda = d + 2 e); columns 2c - 1 and
// 2c = sqrt(2 / n) cos / sin(2 pi i c / n) (lambda = d + 2 e cos(2 pi c / n)); for
// even n the last column is (-1)^i / sqrt(n) (lambda = d - 2 e).  Arguments are
// reduced exactly in integers before sinpi / cospi.
__global__ void k_eig_tridiag_const(double* __restrict__ Q, double* __restrict__ lam, int64_t n, double d, double e,
                                    int periodic) {
  const int64_t total = n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx % n, k = idx / n;
    double v;
    if (!periodic) {
      const int64_t p = ((i + 1) * (k + 1)) % (2 * (n + 1));
      v = sqrt(2.0 / (double)(n + 1)) * sinpi((double)p / (double)(n + 1));
      if (i == 0) lam[k] = d + 2.0 * e * cospi((double)(k + 1) / (double)(n + 1));
    } else {
      double lk;
      if (k == 0) {
        v = 1.0 / sqrt((double)n);
        lk = d + 2.0 * e;
      } else if ((n % 2 == 0) && k == n - 1) {
        v = (i % 2 == 0 ? 1.0 : -1.0) / sqrt((double)n);
        lk = d - 2.0 * e;
      } else {
        const int64_t c = (k + 1) / 2;
        const int64_t p = (i * c) % n;
        const double x = 2.0 * (double)p / (double)n;
        v = sqrt(2.0 / (double)n) * ((k % 2 == 1) ? cospi(x) : sinpi(x));
        lk = d + 2.0 * e * cospi(2.0 * (double)c / (double)n);
      }
      if (i == 0) lam[k] = lk;
    }
    Q[idx] = v;
  }
}

hipError_t launch_eig_tridiag_const(double* Q, double* lam, int64_t n, double d, double e, int periodic,
                                    hipStream_t st) {
  if (n < 1 || (periodic && n < 3)) return hipErrorInvalidValue;
  const int64_t total = n * n;
  const int grid = (int)std::min<int64_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(k_eig_tridiag_const, dim3(grid), dim3(256), 0, st, Q, lam, n, d, e, periodic);
  return hipGetLastError();
}

// zhat_ij <- Ra zhat_ij + Qb bhat_ij with a = lambda_i + mu_j, bhat = fbar coef
// (fbar n x q, coef q x k): N RK4 steps of the decoupled entries
struct FlowLaunch {
  double* Z;
  int64_t n;
  int k;
  const double *lam, *mu, *fbar, *coef;
  int q;
  double h;
  int64_t N;
};
__device__ __forceinline__ void spectral_flow_body(const FlowLaunch& L, int bid, int nblk) {
  double* __restrict__ Z = L.Z;
  const int64_t n = L.n;
  const int k = L.k, q = L.q;
  const double* __restrict__ lam = L.lam;
  const double* __restrict__ mu = L.mu;
  const double* __restrict__ fbar = L.fbar;
  const double* __restrict__ coef = L.coef;
  const double h = L.h;
  const int64_t N = L.N;
  const int64_t total = n * (int64_t)k;
  for (int64_t idx = bid * (int64_t)blockDim.x + threadIdx.x; idx < total; idx += (int64_t)nblk * blockDim.x) {
    const int64_t i = idx % n;
    const int j = (int)(idx / n);
    double b = 0.0;
    for (int qq = 0; qq < q; ++qq) b += fbar[i + qq * n] * coef[qq + j * q];
    double Ra, Qb;
    rk4_power(h * (lam[i] + mu[j]), h, N, Ra, Qb);
    Z[idx] = Ra * Z[idx] + Qb * b;
  }
}

__global__ void k_spectral_flow(FlowLaunch L) { spectral_flow_body(L, blockIdx.x, gridDim.x); }
__global__ void k_spectral_flow_batch(const unsigned char* __restrict__ tab, const int* __restrict__ start,
                                      int count) {
  const int i = batch_member(start, count, blockIdx.x);
  spectral_flow_body(*reinterpret_cast<const FlowLaunch*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes), blockIdx.x - start[i],
                     start[i + 1] - start[i]);
}

hipError_t launch_spectral_flow(double* Z, int64_t n, int k, const double* lam, const double* mu, const double* fbar,
                                const double* coef, int q, double h, int64_t N, hipStream_t st) {
  if (n < 1 || k < 1 || q < 0 || q > kMaxForce || N < 1) return hipErrorInvalidValue;
  const int64_t total = n * (int64_t)k;
  const int grid = (int)std::min<int64_t>((total + 255) / 256, 4096);
  const FlowLaunch L{Z, n, k, lam, mu, fbar, coef, q, h, N};
  if (record_launch(KID_FLOW, 0, (unsigned)grid, 256, 0, L)) return hipSuccess;
  hipLaunchKernelGGL(k_spectral_flow, dim3(grid), dim3(256), 0, st, L);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Symmetric k x k eigen-decomposition (k <= 128) as the one-sided Jacobi SVD of
// M + sigma I with sigma = max_i sum_j |M_ij| >= spectral radius, so the
// shifted matrix is positive semidefinite: its singular vectors are
// eigenvectors of M and its singular values minus sigma the eigenvalues.
// ----------------------------------------------------------------------------

struct ShiftArgs {
  const double* M;
  int ldm, k;
  double *out, *shift, *ev;
};

// the shifted matrix of one problem (one workgroup)
__device__ __forceinline__ void shift_sym_body(const ShiftArgs& A) {
  __shared__ double s_max[256];
  const int t = threadIdx.x;
  const int k = A.k;
  // symmetrised copy (the Galerkin products are symmetric up to rounding), then
  // sigma = max row sum of |M| over the copy
  for (int e = t; e < k * k; e += blockDim.x) {
    const int i = e % k, jj = e / k;
    A.out[e] = 0.5 * (A.M[i + jj * A.ldm] + A.M[jj + i * A.ldm]);
  }
  __syncthreads();
  // row sums in np parts of the columns (fixed order), then over the parts
  const int np = blockDim.x / k;  // k <= 128: at least 2
  if (t < np * k) {
    const int i = t % k, p = t / k;
    double s = 0.0;
    for (int jj = p; jj < k; jj += np) s += fabs(A.out[i + jj * k]);
    s_max[t] = s;
  }
  __syncthreads();
  double mx = 0.0;
  if (t < k) {
    double s = 0.0;
    for (int p = 0; p < np; ++p) s += s_max[t + p * k];
    mx = s;
  }
  __syncthreads();
  s_max[t] = mx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w) s_max[t] = fmax(s_max[t], s_max[t + w]);
    __syncthreads();
  }
  const double sg = s_max[0];
  for (int i = t; i < k; i += blockDim.x) A.out[i + i * k] += sg;
  if (t == 0) *A.shift = sg;
}

__device__ __forceinline__ void unshift_body(const ShiftArgs& A) {
  for (int i = threadIdx.x; i < A.k; i += blockDim.x) A.ev[i] -= *A.shift;
}

// workgroup b: problem b; the batched forms take one pair per record
struct ShiftPair {
  ShiftArgs A0, A1;
};
__global__ __launch_bounds__(256) void k_shift_sym(ShiftArgs A0, ShiftArgs A1) {
  shift_sym_body(blockIdx.x == 0 ? A0 : A1);
}
__global__ void k_unshift(ShiftArgs A0, ShiftArgs A1) { unshift_body(blockIdx.x == 0 ? A0 : A1); }
__global__ __launch_bounds__(256) void k_shift_sym_batch(const unsigned char* __restrict__ tab,
                                                         const int* __restrict__ start, int count) {
  const int i = batch_member(start, count, blockIdx.x);
  const ShiftPair& S = *reinterpret_cast<const ShiftPair*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes);
  shift_sym_body(blockIdx.x - start[i] == 0 ? S.A0 : S.A1);
}
__global__ void k_unshift_batch(const unsigned char* __restrict__ tab, const int* __restrict__ start, int count) {
  const int i = batch_member(start, count, blockIdx.x);
  const ShiftPair& S = *reinterpret_cast<const ShiftPair*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes);
  unshift_body(blockIdx.x - start[i] == 0 ? S.A0 : S.A1);
}

hipError_t launch_sym_eig(const SymEigArgs* P, int count, hipStream_t st) {
  if (count < 1 || count > 2) return hipErrorInvalidValue;
  ShiftArgs sa[2];
  SvdArgs sv[2];
  for (int b = 0; b < count; ++b) {
    const SymEigArgs& p = P[b];
    if (p.k < 1 || p.k > kMaxAug) return hipErrorInvalidValue;
    sa[b] = ShiftArgs{p.M, p.ldm, p.k, p.work, p.work + (size_t)p.k * p.k, p.ev};
    sv[b].S = p.work;
    sv[b].lds = p.k;
    sv[b].k = p.k;
    // the eigenvectors are the orthonormal factor: the accumulated rotations,
    // which the QR-preconditioned sweeps return as the left singular vectors
    // (the right ones have zero columns where M + sigma I is singular)
    sv[b].U = p.W;
    sv[b].V = p.Uscr;
    sv[b].sigma = p.ev;
    sv[b].r_out = p.r_scr;
    sv[b].tol = 0.0;
    sv[b].rmax = p.k;
    sv[b].work = p.svd_work;
  }
  if (count == 1) {
    sa[1] = sa[0];
    sv[1] = sv[0];
  }
  const ShiftPair sp{sa[0], sa[1]};
  if (!record_launch(KID_SHIFT, 0, (unsigned)count, 256, 0, sp))
    hipLaunchKernelGGL(k_shift_sym, dim3(count), dim3(256), 0, st, sa[0], sa[1]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = count == 1 ? launch_jacobi_svd(sv[0], st) : launch_jacobi_svd2(sv[0], sv[1], st);
  if (e != hipSuccess) return e;
  if (!record_launch(KID_UNSHIFT, 0, (unsigned)count, 128, 0, sp))
    hipLaunchKernelGGL(k_unshift, dim3(count), dim3(128), 0, st, sa[0], sa[1]);
  return hipGetLastError();
}

hipError_t launch_de_batched(int kid, int variant, const void* tab, const int* start, int count, unsigned total,
                             unsigned block, size_t lds, hipStream_t st) {
  const unsigned char* t = static_cast<const unsigned char*>(tab);
  const dim3 g(total), b(block);
  switch (kid) {
    case KID_DE_PREP:
      if (variant) hipLaunchKernelGGL(k_de_prep_batch<true>, g, b, lds, st, t, start, count);
      else hipLaunchKernelGGL(k_de_prep_batch<false>, g, b, lds, st, t, start, count);
      break;
    case KID_SSTEP:
      if (variant) hipLaunchKernelGGL(k_s_substep_batch<true>, g, b, lds, st, t, start, count);
      else hipLaunchKernelGGL(k_s_substep_batch<false>, g, b, lds, st, t, start, count);
      break;
    case KID_FLOW: hipLaunchKernelGGL(k_spectral_flow_batch, g, b, lds, st, t, start, count); break;
    case KID_SHIFT: hipLaunchKernelGGL(k_shift_sym_batch, g, b, lds, st, t, start, count); break;
    case KID_UNSHIFT: hipLaunchKernelGGL(k_unshift_batch, g, b, lds, st, t, start, count); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace lri
