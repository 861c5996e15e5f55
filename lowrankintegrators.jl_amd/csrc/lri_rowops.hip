// FP64 kernels on the thin n x r (m x r) factors and the r x r cores.
//
// k_rows is row-parallel over the tall dimension (one thread per row, tiles of
// up to 256 rows, grid-stride over tiles) and may end in a global reduction
// C = out' * W (or the Gram out' * out).  The reduction is finished by the last
// workgroup to arrive, which sums the per-workgroup partials in workgroup order
// (so results are bitwise reproducible) and runs a small epilogue on C in LDS:
// a Cholesky factorisation for CholeskyQR2, the S-step update, the Galerkin
// matrices of the BUG steps.
//
// These replace the LAPACK/BLAS calls of the reference on the factors:
// qr!/Matrix(Q) (projector_splitting.jl:137-138,149-150,174-175,186-187;
// unconventional.jl:141-152; rank_adaptive_unconventional.jl:204-217), the
// U*S / V*S' products (projector_splitting.jl:133,145,170,182), and the r x r
// S-step algebra (projector_splitting.jl:141-142; unconventional.jl:154-156;
// rank_adaptive_unconventional.jl:219-231).
#include "lri_common.h"
#include "lri_kernels.h"

namespace lri {

constexpr int kThreads = 256;
constexpr double kEps = 2.220446049250313e-16;

// ----------------------------------------------------------------------------
// small dense routines on LDS-resident matrices (column-major, ld = k)
// ----------------------------------------------------------------------------

// Upper Cholesky G = R'R of the k x k symmetric matrix G (upper triangle read,
// G is overwritten).  A pivot <= tau = 10 eps max_i G(i,i) marks column p as
// numerically dependent: dropped[p] = 1, R(p,p) = 1 and row p of R is zero, so
// that the inverse computed by triu_inv_drop gives a zero column p in Q.
// Returns the number of dropped columns; sets *nonfinite on NaN/Inf.
__device__ int chol_drop(double* G, double* R, int k, unsigned char* dropped, int* nonfinite) {
  __shared__ double s_tau, s_piv;
  __shared__ int s_nd, s_nf, s_keep;
  const int t = threadIdx.x;
  if (t == 0) {
    double md = 0.0;
    int nf = 0;
    for (int i = 0; i < k; ++i) {
      double d = G[i + i * k];
      if (!isfinite(d)) nf = 1;
      md = fmax(md, d);
    }
    s_tau = 10.0 * kEps * md;
    s_nd = 0;
    s_nf = nf;
  }
  for (int e = t; e < k * k; e += blockDim.x) R[e] = 0.0;
  __syncthreads();
  for (int p = 0; p < k; ++p) {
    if (t == 0) {
      double d = G[p + p * k];
      int keep = (d > s_tau) && isfinite(d);
      if (!isfinite(d)) s_nf = 1;
      s_keep = keep;
      s_piv = keep ? sqrt(d) : 1.0;
      s_nd += keep ? 0 : 1;
      dropped[p] = keep ? 0 : 1;
      R[p + p * k] = s_piv;
    }
    __syncthreads();
    const double piv = s_piv;
    const int keep = s_keep;
    for (int j = p + 1 + t; j < k; j += blockDim.x) R[p + j * k] = keep ? G[p + j * k] / piv : 0.0;
    __syncthreads();
    if (keep) {
      const int rem = k - p - 1;
      for (int e = t; e < rem * rem; e += blockDim.x) {
        int i = p + 1 + e % rem, j = p + 1 + e / rem;
        if (i <= j) G[i + j * k] -= R[p + i * k] * R[p + j * k];
      }
    }
    __syncthreads();
  }
  if (t == 0 && s_nf) *nonfinite = 1;
  __syncthreads();
  return s_nd;
}

// Rinv = R^{-1} for upper-triangular R, with dropped columns/rows zeroed.
__device__ void triu_inv_drop(const double* R, double* Rinv, int k, const unsigned char* dropped) {
  const int t = threadIdx.x;
  for (int e = t; e < k * k; e += blockDim.x) Rinv[e] = 0.0;
  __syncthreads();
  for (int j = t; j < k; j += blockDim.x) {
    if (dropped[j]) continue;
    for (int i = j; i >= 0; --i) {
      if (dropped[i]) continue;
      double s = (i == j) ? 1.0 : 0.0;
      for (int l = i + 1; l <= j; ++l) s -= R[i + l * k] * Rinv[l + j * k];
      Rinv[i + j * k] = s / R[i + i * k];
    }
  }
  __syncthreads();
}

// C = op(A) * op(B), k x k, column-major.
__device__ void small_gemm(const double* A, int tA, const double* B, int tB, double* C, int k) {
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
    int i = e % k, j = e / k;
    double s = 0.0;
    for (int l = 0; l < k; ++l) s += (tA ? A[l + i * k] : A[i + l * k]) * (tB ? B[j + l * k] : B[l + j * k]);
    C[e] = s;
  }
}

// LDS doubles the epilogue needs beyond C (a x b).
__host__ __device__ inline int epi_scratch_doubles(int kind, int a, int b) {
  int k = a > b ? a : b;
  switch (kind) {
    case EPI_CHOL1:
    case EPI_CHOL2:
    case EPI_BUG_N:
      return 2 * k * k;
    case EPI_RA_N:
      return 4 * k * k;
    default:
      return 0;
  }
}

// ----------------------------------------------------------------------------
// epilogues run by the last workgroup on the reduced C (LDS, a x b)
// ----------------------------------------------------------------------------

__device__ void run_epilogue(const EpiArgs& E, double* C, int a, int b, double* scr) {
  const int t = threadIdx.x;
  SmallState* st = E.small;
  switch (E.kind) {
    case EPI_CHOL1:
    case EPI_CHOL2: {
      const int k = a;
      __shared__ unsigned char dropped[kMaxAug];
      double* R = scr;
      double* Rinv = scr + k * k;
      int nf = 0;
      int nd = chol_drop(C, R, k, dropped, &nf);
      triu_inv_drop(R, Rinv, k, dropped);
      if (E.kind == EPI_CHOL1) {
        for (int e = t; e < k * k; e += blockDim.x) {
          st->R1[e] = R[e];
          st->R1inv[e] = Rinv[e];
        }
      } else {
        for (int e = t; e < k * k; e += blockDim.x) st->R2inv[e] = Rinv[e];
        // Rtot = R2 * R1 (R1 was written by an earlier launch on this stream)
        small_gemm(R, 0, st->R1, 0, st->Rtot, k);
        if (E.S) {
          __syncthreads();
          for (int e = t; e < k * k; e += blockDim.x) {
            int i = e % k, j = e / k;
            E.S[i + j * E.ld_s] = E.transpose ? st->Rtot[j + i * k] : st->Rtot[e];
          }
        }
      }
      if (t == 0) {
        if (nf) atomicOr(E.flags, FLAG_NONFINITE);
        if (nd && E.breakdown_is_error) atomicOr(E.flags, FLAG_BREAKDOWN);
        st->ndropped = nd;
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
      const int k = a;
      double* S0 = scr;
      double* T = scr + k * k;
      for (int e = t; e < k * k; e += blockDim.x) {
        st->N[e] = C[e];
        S0[e] = E.S[(e % k) + (e / k) * E.ld_s];
      }
      __syncthreads();
      small_gemm(st->M, 0, S0, 0, T, k);
      __syncthreads();
      small_gemm(T, 0, C, 1, st->Sinit, k);
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
    case EPI_STORE: {
      for (int e = t; e < a * b; e += blockDim.x) st->Ctmp[e] = C[e];
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

// ----------------------------------------------------------------------------
// per-workgroup reduction C_blk += tile_out' * tile_w, register-blocked 4 x 4
// ----------------------------------------------------------------------------

struct RedPlan {
  int a4, b4, nblk, groups, per;
};

__host__ __device__ inline RedPlan red_plan(int a, int b) {
  RedPlan p;
  p.a4 = (a + 3) & ~3;
  p.b4 = (b + 3) & ~3;
  p.nblk = (p.a4 / 4) * (p.b4 / 4);
  if (p.nblk >= kThreads) {
    p.groups = 1;
    p.per = (p.nblk + kThreads - 1) / kThreads;
  } else {
    p.groups = kThreads / p.nblk;
    p.per = 1;
  }
  return p;
}

template <int PER>
__device__ __forceinline__ void red_accum(const RedPlan& p, const double* ta, const double* tb, int nrows,
                                          double (&acc)[PER][16]) {
  const int t = threadIdx.x;
  const int na = p.a4 / 4;
  if (p.groups > 1) {
    const int blk = t % p.nblk, g = t / p.nblk;
    if (g >= p.groups) return;
    const int bi = (blk % na) * 4, bj = (blk / na) * 4;
    for (int row = g; row < nrows; row += p.groups) {
      double x[4], y[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        x[u] = ta[row * p.a4 + bi + u];
        y[u] = tb[row * p.b4 + bj + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[0][u * 4 + v] += x[u] * y[v];
    }
  } else {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int blk = t + q * kThreads;
      if (blk < p.nblk) {
        const int bi = (blk % na) * 4, bj = (blk / na) * 4;
        for (int row = 0; row < nrows; ++row) {
          double x[4], y[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            x[u] = ta[row * p.a4 + bi + u];
            y[u] = tb[row * p.b4 + bj + u];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[q][u * 4 + v] += x[u] * y[v];
        }
      }
    }
  }
}

template <int PER>
__device__ void red_flush(const RedPlan& p, int a, int b, double (&acc)[PER][16], double* red, double* gpart) {
  const int t = threadIdx.x;
  const int na = p.a4 / 4;
  if (p.groups > 1) {
    const int blk = t % p.nblk, g = t / p.nblk;
    const int bi = (blk % na) * 4, bj = (blk / na) * 4;
    if (g < p.groups)
      for (int u = 0; u < 4; ++u)
        for (int v = 0; v < 4; ++v) red[g * p.a4 * p.b4 + (bi + u) + (bj + v) * p.a4] = acc[0][u * 4 + v];
    __syncthreads();
    for (int e = t; e < a * b; e += kThreads) {
      int i = e % a, j = e / a;
      double s = 0.0;
      for (int gg = 0; gg < p.groups; ++gg) s += red[gg * p.a4 * p.b4 + i + j * p.a4];
      gpart[e] = s;
    }
  } else {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int blk = t + q * kThreads;
      if (blk < p.nblk) {
        const int bi = (blk % na) * 4, bj = (blk / na) * 4;
        for (int u = 0; u < 4; ++u)
          for (int v = 0; v < 4; ++v) {
            int i = bi + u, j = bj + v;
            if (i < a && j < b) gpart[i + j * a] = acc[q][u * 4 + v];
          }
      }
    }
  }
}

// The last-arriving workgroup sums the partials (workgroup order) into C and
// runs the epilogue; it then re-arms the counter for the next launch.
__device__ void finish_reduction(const EpiArgs& E, int a, int b, const double* gpart_all, double* C,
                                 double* scr) {
  __shared__ int s_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned prev = atomicAdd(E.counter, 1u);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  const int nb = gridDim.x;
  for (int e = threadIdx.x; e < a * b; e += kThreads) {
    double s = 0.0;
    for (int blk = 0; blk < nb; ++blk) s += __builtin_nontemporal_load(&gpart_all[(int64_t)blk * a * b + e]);
    C[e] = s;
  }
  __syncthreads();
  run_epilogue(E, C, a, b, scr);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    atomicExch(E.counter, 0u);
  }
}

// ----------------------------------------------------------------------------
// the row kernel
//   out[i, c] = (parts ? sum_s part[s][c][i] : 0) + base_sign * base[i, c]
//             + xy_sign * sum_d X[i, d] * op(Y)(d, c)
// LDS: Y (r_in x r_out) | tile_out (tr x a4) | tile_w (tr x b4) | group partials,
//      and after the tile loop, C (a x b) + epilogue scratch in the tile area.
// ----------------------------------------------------------------------------

template <typename PT, int PER>
__global__ __launch_bounds__(kThreads) void k_rows(RowArgs<PT> A) {
  extern __shared__ double lds[];
  const int t = threadIdx.x;
  const int ro = A.r_out, ri = A.r_in;
  const int rw = A.W ? A.rw : ro;
  double* Ys = lds;
  double* area = Ys + A.ri_pad * A.ro_pad;
  const RedPlan p = red_plan(ro, rw);
  double* tout = area;
  double* tw = tout + A.tr * p.a4;
  double* red = tw + (A.W ? A.tr * p.b4 : 0);

  if (A.X) {
    for (int e = t; e < ri * ro; e += kThreads) {
      int d = e % ri, c = e / ri;
      Ys[d + c * ri] = A.transY ? A.Y[c + d * A.ldy] : A.Y[d + c * A.ldy];
    }
  }
  __syncthreads();

  double acc[PER][16];
#pragma unroll
  for (int q = 0; q < PER; ++q)
#pragma unroll
    for (int u = 0; u < 16; ++u) acc[q][u] = 0.0;

  const int64_t ntiles = (A.rows + A.tr - 1) / A.tr;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t row0 = tile * A.tr;
    const int nrows = (int)min((int64_t)A.tr, A.rows - row0);
    if (t < A.tr) {
      const int64_t i = row0 + t;
      const bool ok = t < nrows;
      for (int c = 0; c < ro; ++c) {
        double v = 0.0;
        if (ok) {
          if (A.parts) {
            for (int s = 0; s < A.nsplit; ++s) v += (double)A.parts[((int64_t)s * A.rb + c) * A.rows + i];
            if (A.sum_out) A.sum_out[i + c * A.ld_sum] = v;
          }
          if (A.base) v += A.base_sign * A.base[i + c * A.ld_base];
          if (A.X) {
            double s = 0.0;
            for (int d = 0; d < ri; ++d) s += A.X[i + d * A.ldx] * Ys[d + c * ri];
            v += A.xy_sign * s;
          }
          if (A.out) A.out[i + c * A.ldo] = v;
          if (A.out2) A.out2[i + c * A.ldo2] = v;
          if (A.pk_f64) A.pk_f64[i * A.pk_cols + c] = v;
          if (A.pk_f32) A.pk_f32[i * A.pk_cols + c] = (float)v;
          if (A.pk_hi) {
            uint32_t hi, lo;
            split_bf16((float)v, hi, lo);
            A.pk_hi[c * A.ld_pk + i] = (uint16_t)hi;
            A.pk_lo[c * A.ld_pk + i] = (uint16_t)lo;
          }
        }
        if (A.reduce) tout[t * p.a4 + c] = ok ? v : 0.0;
      }
      if (ok) {  // zero the packed columns beyond the current rank
        for (int c = ro; c < A.pk_cols; ++c) {
          if (A.pk_f64) A.pk_f64[i * A.pk_cols + c] = 0.0;
          if (A.pk_f32) A.pk_f32[i * A.pk_cols + c] = 0.0f;
          if (A.pk_hi) {
            A.pk_hi[c * A.ld_pk + i] = 0;
            A.pk_lo[c * A.ld_pk + i] = 0;
          }
        }
      }
      if (A.reduce) {
        for (int c = ro; c < p.a4; ++c) tout[t * p.a4 + c] = 0.0;
        if (A.W)
          for (int c = 0; c < p.b4; ++c) tw[t * p.b4 + c] = (ok && c < rw) ? A.W[i + c * A.ldw] : 0.0;
      }
    }
    if (A.reduce) {
      __syncthreads();
      red_accum<PER>(p, tout, A.W ? tw : tout, nrows, acc);
      __syncthreads();
    }
  }
  if (!A.reduce) return;
  double* gpart = A.gpart + (int64_t)blockIdx.x * ro * rw;
  red_flush<PER>(p, ro, rw, acc, red, gpart);
  __syncthreads();
  double* C = area;
  double* scr = area + ro * rw;
  finish_reduction(A.epi, ro, rw, A.gpart, C, scr);
}

template <typename PT>
hipError_t launch_rows(RowArgs<PT> A, hipStream_t st) {
  const int rw = A.W ? A.rw : A.r_out;
  const RedPlan p = red_plan(A.r_out, rw);
  const int width = p.a4 + (A.W ? p.b4 : 0);
  int tr = 256;
  while (tr > 16 && (size_t)tr * width * 8 > 65536) tr /= 2;
  A.tr = tr;
  A.ri_pad = A.X ? A.r_in : 0;
  A.ro_pad = A.X ? A.r_out : 0;
  size_t body = (size_t)tr * width + (p.groups > 1 ? (size_t)p.groups * p.a4 * p.b4 : 0);
  size_t epi = A.reduce ? (size_t)A.r_out * rw + epi_scratch_doubles(A.epi.kind, A.r_out, rw) : 0;
  size_t lds = sizeof(double) * ((size_t)A.ri_pad * A.ro_pad + (body > epi ? body : epi));
  if (lds > 150 * 1024) return hipErrorInvalidValue;  // leaves room for the static shared variables
  const int64_t ntiles = (A.rows + tr - 1) / tr;
  int grid = (int)(ntiles < A.max_blocks ? ntiles : A.max_blocks);
  if (grid < 1) grid = 1;
  if (p.per <= 1) {
    hipLaunchKernelGGL((k_rows<PT, 1>), dim3(grid), dim3(kThreads), lds, st, A);
  } else if (p.per <= 4) {
    hipLaunchKernelGGL((k_rows<PT, 4>), dim3(grid), dim3(kThreads), lds, st, A);
  } else {
    return hipErrorInvalidValue;
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

__global__ __launch_bounds__(kThreads) void k_jacobi_svd(SvdArgs P) {
  extern __shared__ double lds[];
  const int k = P.k;
  const int ke = (k + 1) & ~1;
  double* A = lds;                   // k rows x ke columns
  double* sig = A + (size_t)ke * k;  // ke
  double* Vm = P.work ? P.work : sig + ke;  // ke x ke, LDS when it fits
  const int t = threadIdx.x;
  for (int e = t; e < ke * k; e += kThreads) {
    int i = e % k, j = e / k;
    A[e] = (j < k) ? P.S[i + j * P.lds] : 0.0;
  }
  for (int e = t; e < ke * ke; e += kThreads) Vm[e] = ((e % ke) == (e / ke)) ? 1.0 : 0.0;
  __syncthreads();
  const int npairs = ke / 2;
  int gsz = 1;
  while (gsz * 2 * npairs <= kThreads && gsz < 64) gsz *= 2;
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
          double x = A[i + pa * k], y = A[i + pb * k];
          al += x * x;
          be += y * y;
          ga += x * y;
        }
      }
      for (int off = gsz / 2; off > 0; off >>= 1) {
        al += __shfl_xor(al, off, gsz);
        be += __shfl_xor(be, off, gsz);
        ga += __shfl_xor(ga, off, gsz);
      }
      if (pa >= 0 && ga != 0.0 && fabs(ga) > 4.0 * kEps * sqrt(al * be)) {
        double zeta = (be - al) / (2.0 * ga);
        double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double c = 1.0 / sqrt(1.0 + tt * tt);
        double s = c * tt;
        for (int i = ln; i < k; i += gsz) {
          double x = A[i + pa * k], y = A[i + pb * k];
          A[i + pa * k] = c * x - s * y;
          A[i + pb * k] = s * x + c * y;
        }
        for (int i = ln; i < ke; i += gsz) {
          double x = Vm[i + pa * ke], y = Vm[i + pb * ke];
          Vm[i + pa * ke] = c * x - s * y;
          Vm[i + pb * ke] = s * x + c * y;
        }
        if (ln == 0) s_rot = 1;
      }
      __syncthreads();
    }
    const int any = s_rot;
    __syncthreads();
    if (!any) break;
  }
  for (int j = t; j < ke; j += kThreads) {
    double s = 0.0;
    if (j < k)
      for (int i = 0; i < k; ++i) s += A[i + j * k] * A[i + j * k];
    sig[j] = (j < k) ? sqrt(s) : -1.0;
  }
  __syncthreads();
  for (int j = t; j < k; j += kThreads) {
    int pos = 0;
    for (int l = 0; l < k; ++l) pos += (sig[l] > sig[j]) || (sig[l] == sig[j] && l < j);
    P.sigma[pos] = sig[j];
    const double inv = sig[j] > 0.0 ? 1.0 / sig[j] : 0.0;
    for (int i = 0; i < k; ++i) {
      P.U[i + pos * k] = A[i + j * k] * inv;
      P.V[i + pos * k] = Vm[i + j * ke];
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

hipError_t launch_jacobi_svd(SvdArgs P, hipStream_t st) {
  if (P.k < 1 || P.k > kMaxAug) return hipErrorInvalidValue;
  const int ke = (P.k + 1) & ~1;
  size_t lds = sizeof(double) * ((size_t)ke * P.k + ke);
  if (lds + sizeof(double) * (size_t)ke * ke <= 150 * 1024) {
    lds += sizeof(double) * (size_t)ke * ke;
    P.work = nullptr;
  } else if (!P.work) {
    return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL(k_jacobi_svd, dim3(1), dim3(kThreads), lds, st, P);
  return hipGetLastError();
}

}  // namespace lri
