// Context, device buffers and step orchestration behind the C ABI (include/lri.h).
//
// One context = one problem size (n x m, rank capacity rcap) on one GPU, one
// HIP stream.  All buffers are allocated once at lri_create; a step is a fixed
// sequence of stream-ordered launches with no host synchronisation (except the
// rank read-back of the rank-adaptive step).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/lri.h"
#include "lri_common.h"
#include "lri_kernels.h"

using namespace lri;

namespace {

thread_local std::string g_create_error;

struct Failure {
  int status;
  std::string msg;
};

#define HIPCHECK(call)                                                                   \
  do {                                                                                   \
    hipError_t _e = (call);                                                              \
    if (_e != hipSuccess)                                                                \
      throw Failure{_e == hipErrorOutOfMemory ? LRI_EOOM : LRI_EHIP,                     \
                    std::string(#call) + ": " + hipGetErrorString(_e)};                  \
  } while (0)

#define REQUIRE(cond, msg)                                   \
  do {                                                       \
    if (!(cond)) throw Failure{LRI_EINVAL, std::string(msg)}; \
  } while (0)

// S(0:r, 0:r) = S(0:r, 0:r)' in place (ld lds); one workgroup, each pair swapped by one thread
__global__ void k_transpose_sq(double* S, int lds, int r) {
  for (int e = threadIdx.x; e < r * r; e += blockDim.x) {
    const int i = e % r, j = e / r;
    if (i < j) {
      const double a = S[i + j * lds], b = S[j + i * lds];
      S[i + j * lds] = b;
      S[j + i * lds] = a;
    }
  }
}

// A = I (k x k, ld k)
__global__ void k_eye(double* A, int k) {
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) A[e] = (e % k == e / k) ? 1.0 : 0.0;
}

// S (rcap x rcap, ld lds) = diag(sigma[0..r-1], 0, ...); one workgroup
struct SetDiagArgs {
  double* S;
  const double* sigma;
  int lds, rcap, r;
};
__device__ __forceinline__ void set_diag_body(const SetDiagArgs& a) {
  for (int e = threadIdx.x; e < a.rcap * a.rcap; e += blockDim.x) {
    int i = e % a.rcap, j = e / a.rcap;
    a.S[i + j * a.lds] = (i == j && i < a.r) ? a.sigma[i] : 0.0;
  }
}
__global__ void k_set_diag(SetDiagArgs a) { set_diag_body(a); }

// dst[0..n) = src[0..n) (FP64 words), the workgroups striding over the range
struct CopyArgs {
  double* dst;
  const double* src;
  int64_t n;
};
constexpr int kCopyThreads = 256;
__device__ __forceinline__ void copy_body(const CopyArgs& a, int bid, int nblk) {
  for (int64_t i = (int64_t)bid * kCopyThreads + threadIdx.x; i < a.n; i += (int64_t)nblk * kCopyThreads)
    a.dst[i] = a.src[i];
}

// the batched forms (argument blocks and workgroup offsets as in lri_kernels.h)
__global__ __launch_bounds__(256) void k_set_diag_batch(const unsigned char* __restrict__ tab,
                                                        const int* __restrict__ start, int count) {
  const int i = batch_member(start, count, blockIdx.x);
  set_diag_body(*reinterpret_cast<const SetDiagArgs*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes));
}
__global__ __launch_bounds__(kCopyThreads) void k_copy_batch(const unsigned char* __restrict__ tab,
                                                             const int* __restrict__ start, int count) {
  const int i = batch_member(start, count, blockIdx.x);
  copy_body(*reinterpret_cast<const CopyArgs*>(tab + (size_t)start[count + 1 + i] * kRecArgBytes),
            blockIdx.x - start[i], start[i + 1] - start[i]);
}

// Per-member status words of a batch: out[i] = the new rank of member i's
// rank-adaptive step, out[count + i] = its condition flags, which are cleared.
struct MemberWords {
  const int* r_new;
  int* flags;
};
__global__ __launch_bounds__(256) void k_gather_status(const MemberWords* __restrict__ w, int* __restrict__ out,
                                                       int count) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  out[i] = *w[i].r_new;
  out[count + i] = *w[i].flags;
  *w[i].flags = 0;
}

struct Timed {
  int family;
  hipEvent_t a, b;
};

// small FP64 matrices of a matrix-ODE step, device-resident (k <= 2 rcap)
struct DeSmall {
  double Ga[kMaxAug * kMaxAug];  // U'AU (n side)
  double Gb[kMaxAug * kMaxAug];  // V'BV (m side)
  double Wa[kMaxAug * kMaxAug];  // eigenvectors of Ga
  double Wb[kMaxAug * kMaxAug];
  double alpha[kMaxAug];         // eigenvalues of Ga
  double beta[kMaxAug];
  double gaT[kMaxForce * kMaxAug];  // G'U (q x k)
  double hbT[kMaxForce * kMaxAug];  // H'V (q x k)
  double CW[2][kMaxAug * kMaxAug];       // op(C0) W of up to two concurrent sub-problems
  double coef[2][kMaxForce * kMaxAug];
  double Snew[kMaxAug * kMaxAug];
  double ework[2][kMaxAug * kMaxAug + 2];  // shifted matrix + shift, per concurrent eigenproblem
  double Uscr[2][kMaxAug * kMaxAug];
  double svd2[kSvdWorkDoubles];            // global Jacobi work of the second eigenproblem (k > 64)
  double swork[2 * kMaxAug * kMaxAug + 2 * kMaxAug * kMaxForce];
  int rscr[2];
};

// Galerkin data, tensors and cotangents of the matrix-ODE BUG adjoint
// (device-resident; k <= kQuadMaxRank, ld k; index 0: the step's start bases
// U0, V0, index 1: the new bases U1, V1)
constexpr int kQ2 = kQuadMaxRank * kQuadMaxRank;
constexpr int kQ3 = kQ2 * kQuadMaxRank;
constexpr int kQPart = kQ3 + kQ2 + kMaxForce * kQuadMaxRank + 1;
struct QuadAdj {
  double Ga0[kQ2], Gb0[kQ2], gaT0[kMaxForce * kQuadMaxRank], hbT0[kMaxForce * kQuadMaxRank];
  double Gab1[kQ2], Gbb1[kQ2], gaTb1[kMaxForce * kQuadMaxRank], hbTb1[kMaxForce * kQuadMaxRank];
  double Sib[kQ2];  // cotangent of the S sub-step's start M S0 N'
  double scal[2];   // derivatives along the scales of A and B
  double W0[kQ3], T0[kQ3], Wb1[kQ3], Tb1[kQ3];
  // summed reverse-stage partials: K side [T0bar | Gb0bar | hbT0bar | <kbar, A X>],
  // L side [W0bar | Ga0bar | gaT0bar | <lbar, B X>]
  double sumK[kQPart], sumL[kQPart];
};

}  // namespace

struct lri_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t n = 0, m = 0, rcap = 0, r = 0;
  int prec = LRI_PREC_F64;
  // FP64 state and work (column-major)
  double *U = nullptr, *V = nullptr, *S = nullptr;  // n x rcap, m x rcap, rcap x rcap
  double *U2 = nullptr, *V2 = nullptr;              // BUG: new bases before the swap
  double *K = nullptr, *K2 = nullptr, *P = nullptr;  // n x 2rcap
  double *L = nullptr, *L2 = nullptr, *Q = nullptr;  // m x 2rcap
  double *Uhat = nullptr, *Vhat = nullptr;           // n x 2rcap, m x 2rcap
  // packed B operands: N side (rows m) and T side (rows n)
  double *pkN_f64 = nullptr, *pkT_f64 = nullptr;
  float *pkN_f32 = nullptr, *pkT_f32 = nullptr;
  uint16_t *pkN_hi = nullptr, *pkN_lo = nullptr, *pkT_hi = nullptr, *pkT_lo = nullptr;
  int64_t ldm_pk = 0, ldn_pk = 0;
  // packed-panel validity: which pass mode and rb they hold (-1: stale)
  int pkN_mode = -1, pkN_rb = 0, pkT_mode = -1, pkT_rb = 0;
  // split partials
  void *partN = nullptr, *partT = nullptr;
  size_t part_bytes = 0;
  // reductions
  double* gpart = nullptr;
  int* flags = nullptr;
  SmallState* small = nullptr;
  double* svd_work = nullptr;
  // snapshots
  std::vector<const void*> snaps;
  std::vector<void*> owned;
  int sdtype = -1;
  int64_t sld = 0;
  // timing
  bool timing = false;
  bool timing_passes_only = false;
  std::vector<Timed> pending;
  std::vector<hipEvent_t> pool;
  double tot_ms[LRI_KERNEL_COUNT] = {};
  int64_t launches[LRI_KERNEL_COUNT] = {};
  std::vector<float> samples[LRI_KERNEL_COUNT];  // per-launch durations (ms)
  int timing_every = 1;                           // time one launch in `timing_every` per family
  int64_t seen[LRI_KERNEL_COUNT] = {};
  std::string err;
  // matrix-ODE operator F(Y) = A Y + Y B' + G H' (allocated on first use)
  bool op_set = false;
  double *op_ad = nullptr, *op_ae = nullptr, *op_bd = nullptr, *op_be = nullptr;
  double *op_G = nullptr, *op_H = nullptr;
  int op_aper = 0, op_bper = 0, op_q = 0;
  DeSmall* de = nullptr;
  bool pack_enabled = true;
  // Transposed view (swap_sides): the context seen as the transposed problem,
  // n <-> m, U <-> V and every side's buffers exchanged, the data A' in place of
  // A: a pass over dA in the view is the other direction's pass over the real
  // snapshots.  The dual KSL step is the primal step of the transposed problem,
  // so its adjoint runs the primal adjoint in this view.
  bool tview = false;
  void swap_sides() {
    std::swap(n, m);
    std::swap(U, V);
    std::swap(U2, V2);
    std::swap(K, L);
    std::swap(K2, L2);
    std::swap(P, Q);
    std::swap(Uhat, Vhat);
    std::swap(pkN_f64, pkT_f64);
    std::swap(pkN_f32, pkT_f32);
    std::swap(pkN_hi, pkT_hi);
    std::swap(pkN_lo, pkT_lo);
    std::swap(ldm_pk, ldn_pk);
    std::swap(pkN_mode, pkT_mode);
    std::swap(pkN_rb, pkT_rb);
    std::swap(partN, partT);
    tview = !tview;
  }
  void transpose_S() {
    hipLaunchKernelGGL(k_transpose_sq, dim3(1), dim3(256), 0, stream, S, (int)rcap, (int)r);
    HIPCHECK(hipGetLastError());
  }
  // constant-coefficient operators: analytic eigenbasis Q (rows x rows), its
  // eigenvalues, and Q' G / Q' H, for the closed-form RK4 sub-steps (one per side;
  // side 1 may alias side 0)
  bool spec[2] = {false, false};
  double *specQ[2] = {nullptr, nullptr}, *specLam[2] = {nullptr, nullptr}, *specF[2] = {nullptr, nullptr};
  bool spec_alias = false;
  // Resident eigenbasis (both sides constant-coefficient, linear operator): the
  // device state holds U^ = Qa' U and V^ = Qb' V, and the operator is its
  // diagonal form (A -> diag(lam_a), B -> diag(lam_b), G -> Qa' G, H -> Qb' H),
  // so a step needs no n x n transforms; the state is transformed at
  // lri_set_state / lri_get_state.  res_saved keeps the operator's own arrays.
  bool resident = false;
  double* res_zero = nullptr;  // max(n, m) zeros: the off-diagonals of the diagonal forms
  double* res_saved[6] = {};   // op_ad, op_ae, op_bd, op_be, op_G, op_H
  int res_saved_per[2] = {0, 0};
  std::vector<char> traj_resident;  // slot saved while resident
  // The Galerkin matrices U'AU, V'BV, U'G, H'V and their eigendecompositions in
  // the DE work area belong to the current U and V: a BUG step ends with the
  // S sub-step's ones, which are the next BUG step's K/L ones. Any other entry
  // point clears the flag.
  bool de_gal_fresh = false;
  int* r_host = nullptr;  // pinned: the rank of a rank-adaptive step
  int ra_k2 = 0;
  bool defer_rank_copy = false;
  BugAdjSmall* adj = nullptr;  // r x r work of the BUG adjoint (allocated on first use)
  RaAdjSmall* radj = nullptr;  // 2r x 2r work of the rank-adaptive adjoint (allocated on first use)
  // captured rank-adaptive DE steps (up to the SVD), one per rank and step parameters
  struct RaGraph {
    int64_t r, nsub, rmax;
    double dt, tol;
    hipGraphExec_t exec;
  };
  std::vector<RaGraph> ra_graphs;
  // bumped whenever recorded or captured launches of this context go stale (a
  // new operator or basis): the merged launches of a batch key on it
  uint64_t epoch = 0;
  // the context's own stream; a batch with merged launches points `stream` at
  // its shared stream instead
  hipStream_t own_stream = nullptr;
  // device-resident trajectory (the deep copies of update_sol!, primitives.jl:82-90):
  // slot k holds U (n x rcap), S (rcap x rcap) and V (m x rcap) of one saved state
  double* traj = nullptr;
  int64_t traj_slots = 0;
  std::vector<int64_t> traj_rank;  // rank of each slot (0: never saved)
  int64_t traj_stride() const { return n * rcap + rcap * rcap + m * rcap; }
  // quadratic term -(D Y) o Y (D tridiagonal on the n rows), its tensors
  // T = V x V x V and W = (D U) x U x U, and their scratch
  bool op_quad = false;
  double *op_dl = nullptr, *op_dd = nullptr, *op_du = nullptr;
  int op_dper = 0;
  double *quad_part = nullptr, *quad_T = nullptr, *quad_W = nullptr, *quad_work = nullptr;
  // matrix-ODE adjoint (allocated on first use): r x r data and the row work
  // (sub-step starts of K and L, stage inputs, cotangents, stage partials)
  QuadAdj* qadj = nullptr;
  double* qadj_rows = nullptr;
  size_t qadj_rows_len = 0;

  // ------------------------------------------------------------------ helpers
  int pass_mode() const {
    if (prec == LRI_PREC_F64 || sdtype == LRI_F64) return PASS_F64;
    return sdtype == LRI_BF16 ? PASS_BF16X3 : PASS_F32;
  }
  static int rb_for(int mode, int64_t cols) {
    if (mode == PASS_F64) {
      if (cols <= 16) return 16;
      if (cols <= 32) return 32;
      if (cols <= 48) return 48;
      if (cols <= 64) return 64;
      return 128;
    }
    if (cols <= 32) return 32;
    if (cols <= 64) return 64;
    return 128;
  }

  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }
  template <class F>
  void timed(int family, F&& f) {
    if (!timing || (timing_passes_only && family == LRI_KERNEL_SMALL) || (seen[family]++ % timing_every) != 0) {
      HIPCHECK(f());
      return;
    }
    Timed t{family, ev(), ev()};
    HIPCHECK(hipEventRecord(t.a, stream));
    HIPCHECK(f());
    HIPCHECK(hipEventRecord(t.b, stream));
    pending.push_back(t);
  }
  void drain_timing() {
    for (auto& t : pending) {
      float ms = 0.f;
      HIPCHECK(hipEventSynchronize(t.b));
      HIPCHECK(hipEventElapsedTime(&ms, t.a, t.b));
      tot_ms[t.family] += ms;
      launches[t.family] += 1;
      samples[t.family].push_back(ms);
      pool.push_back(t.a);
      pool.push_back(t.b);
    }
    pending.clear();
  }

  template <typename PT>
  void rows(RowArgs<PT> A) {
    A.epi.small = small;
    A.epi.flags = flags;
    A.pre.small = small;
    A.pre.flags = flags;
    A.gpart = gpart;
    timed(LRI_KERNEL_SMALL, [&] { return launch_rows<PT>(A, stream); });
  }

  // Pack X (rows x cols FP64, ld ldx) as the B operand of a pass.
  void pack(const double* X, int64_t ldx, int64_t rows_, int64_t cols, bool forN) {
    const int mode = pass_mode();
    const int rb = rb_for(mode, cols);
    RowArgs<double> A;
    A.rows = rows_;
    A.r_out = (int)cols;
    A.base = X;
    A.ld_base = ldx;
    set_pack_outputs(A, forN, mode, rb);
    rows(A);
  }
  template <typename PT>
  void set_pack_outputs(RowArgs<PT>& A, bool forN, int mode, int rb) {
    if (!pack_enabled) return;
    A.pk_cols = rb;
    if (mode == PASS_F64) {
      A.pk_f64 = forN ? pkN_f64 : pkT_f64;
    } else if (mode == PASS_F32) {
      A.pk_f32 = forN ? pkN_f32 : pkT_f32;
      A.pk_layout = (forN != tview) ? 0 : 1;  // the layout of the real pass that reads it
    } else {
      A.pk_hi = forN ? pkN_hi : pkT_hi;
      A.pk_lo = forN ? pkN_lo : pkT_lo;
      A.pk_layout = (forN != tview) ? 0 : 1;
    }
    if (forN) {
      pkN_mode = mode;
      pkN_rb = rb;
    } else {
      pkT_mode = mode;
      pkT_rb = rb;
    }
  }
  void ensure_packed(bool forN) {
    const int mode = pass_mode();
    const int rb = rb_for(mode, r);
    if (forN && (pkN_mode != mode || pkN_rb != rb)) pack(V, m, m, r, true);
    if (!forN && (pkT_mode != mode || pkT_rb != rb)) pack(U, n, n, r, false);
  }

  // Reduction chunks chosen by lri_autotune_passes for this context, per pass
  // shape; the other shapes take plan_pass's default.
  struct TunedPass {
    int mode;
    bool transposed;
    int64_t out_rows, red_len;
    int rb;
    int64_t chunk;
  };
  std::vector<TunedPass> tuned;
  int64_t plan(int mode, bool transposed, int64_t out_rows, int64_t red_len, int rb, int* splits, int64_t* chunk) {
    for (const TunedPass& t : tuned)
      if (t.mode == mode && t.transposed == transposed && t.out_rows == out_rows && t.red_len == red_len && t.rb == rb)
        return plan_pass_chunk(out_rows, red_len, rb, t.chunk, splits, chunk);
    return plan_pass(mode, transposed, out_rows, red_len, rb, splits, chunk);
  }

  // The arguments of a pass over dA = A[kn] - A[kp] (kp < 0: over A[kn]
  // itself) with the currently packed B of width `cols` (splits not planned).
  PassArgs pass_args(bool transposed, int64_t kp, int64_t kn, int64_t cols, void* part) {
    const int mode = pass_mode();
    PassArgs a;
    a.anext = snaps[kn];
    a.aprev = kp >= 0 ? snaps[kp] : nullptr;
    a.dtype = sdtype;
    a.lda = sld;
    a.n = tview ? m : n;  // the snapshots' own dimensions
    a.m = tview ? n : m;
    a.rb = rb_for(mode, cols);
    if (!transposed) {
      REQUIRE(pkN_mode == mode && pkN_rb == a.rb, "internal: stale packed V panel");
      if (mode == PASS_F64) a.B = pkN_f64;
      else if (mode == PASS_F32) a.B = pkN_f32;
      else {
        a.B = pkN_hi;
        a.B2 = pkN_lo;
        a.ldb = ldm_pk;
      }
    } else {
      REQUIRE(pkT_mode == mode && pkT_rb == a.rb, "internal: stale packed U panel");
      if (mode == PASS_F64) a.B = pkT_f64;
      else if (mode == PASS_F32) a.B = pkT_f32;
      else {
        a.B = pkT_hi;
        a.B2 = pkT_lo;
        a.ldb = ldn_pk;
      }
    }
    a.part = part;
    return a;
  }
  // The product passes with the currently packed B; they return the split count.
  // In the transposed view pass N over dA' is pass T over the real snapshots.
  int pass_dir(bool transposed, int64_t kp, int64_t kn, int64_t cols, void* part) {
    const int mode = pass_mode();
    PassArgs a = pass_args(transposed, kp, kn, cols, part);
    const bool real_t = transposed != tview;
    int64_t need = real_t ? plan(mode, true, a.m, a.n, a.rb, &a.splits, &a.chunk)
                          : plan(mode, false, a.n, a.m, a.rb, &a.splits, &a.chunk);
    REQUIRE((size_t)need * (mode == PASS_F64 ? 8 : 4) <= part_bytes, "internal: partial buffer too small");
    timed(transposed ? LRI_KERNEL_PASS_T : LRI_KERNEL_PASS_N,
          [&] { return real_t ? launch_passT(mode, a, stream) : launch_passN(mode, a, stream); });
    return a.splits;
  }
  int passN(int64_t kp, int64_t kn, int64_t cols, void* part) { return pass_dir(false, kp, kn, cols, part); }
  int passT(int64_t kp, int64_t kn, int64_t cols, void* part) { return pass_dir(true, kp, kn, cols, part); }

  // The pass autotuner: for pass N and pass T over dA = A[kn] - A[kp] at the
  // current rank and precision, time every candidate reduction chunk
  // (pass_chunk_candidates: about 128 .. 3072 workgroups) over `reps` launches
  // after one untimed launch, and keep the fastest for this context.
  // report[4 d ..] for d = 0 (pass N), 1 (pass T): chunk, splits, mean
  // microseconds of the chosen chunk, mean microseconds of the default chunk.
  void autotune_passes(int64_t kp, int64_t kn, int reps, double* report) {
    const int mode = pass_mode();
    const int rb = rb_for(mode, r);
    const size_t es = mode == PASS_F64 ? 8 : 4;
    ensure_packed(true);
    ensure_packed(false);
    for (int dir = 0; dir < 2; ++dir) {
      const bool tr = dir == 1;
      const int64_t out_rows = tr ? m : n, red_len = tr ? n : m;
      int64_t cand[16];
      const int nc = pass_chunk_candidates(mode, tr, out_rows, red_len, rb, cand, 16);
      size_t need = 0;
      for (int k = 0; k < nc; ++k) {
        int sp;
        int64_t ch;
        need = std::max(need, es * (size_t)plan_pass_chunk(out_rows, red_len, rb, cand[k], &sp, &ch));
      }
      if (need > part_bytes) {  // room for the largest candidate's partials
        HIPCHECK(hipStreamSynchronize(stream));
        HIPCHECK(hipFree(partN));
        HIPCHECK(hipFree(partT));
        partN = partT = nullptr;
        HIPCHECK(hipMalloc(&partN, need));
        HIPCHECK(hipMalloc(&partT, need));
        part_bytes = need;
      }
      tuned.erase(std::remove_if(tuned.begin(), tuned.end(),
                                 [&](const TunedPass& t) {
                                   return t.mode == mode && t.transposed == tr && t.out_rows == out_rows &&
                                          t.red_len == red_len && t.rb == rb;
                                 }),
                  tuned.end());
      int s0;
      int64_t ch0;
      plan_pass(mode, tr, out_rows, red_len, rb, &s0, &ch0);
      PassArgs a = pass_args(tr, kp, kn, r, tr ? partT : partN);
      hipEvent_t e0 = ev(), e1 = ev();
      // three interleaved rounds over the candidates, the median per candidate (a
      // drift of the device's clocks during the sweep then does not favour the
      // candidates timed last); the default plan is kept unless a candidate is
      // at least 1 % faster, so that noise alone does not switch plans
      constexpr int kRounds = 3;
      std::vector<double> t_us((size_t)nc * kRounds);
      std::vector<int64_t> chunk_of((size_t)nc);
      for (int round = 0; round < kRounds; ++round)
        for (int k = 0; k < nc; ++k) {
          plan_pass_chunk(out_rows, red_len, rb, cand[k], &a.splits, &a.chunk);
          chunk_of[(size_t)k] = a.chunk;
          auto run = [&] { HIPCHECK(tr ? launch_passT(mode, a, stream) : launch_passN(mode, a, stream)); };
          run();
          HIPCHECK(hipEventRecord(e0, stream));
          for (int rep = 0; rep < reps; ++rep) run();
          HIPCHECK(hipEventRecord(e1, stream));
          HIPCHECK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
          t_us[(size_t)k * kRounds + round] = 1e3 * ms / reps;
        }
      pool.push_back(e0);
      pool.push_back(e1);
      double best = 0.0, dflt = 0.0;
      int64_t bestc = ch0;
      bool have_dflt = false;
      for (int k = 0; k < nc; ++k) {
        double* tk = t_us.data() + (size_t)k * kRounds;
        std::sort(tk, tk + kRounds);
        const double med = tk[kRounds / 2];
        if (chunk_of[(size_t)k] == ch0) {
          dflt = med;
          have_dflt = true;
        }
        if (k == 0 || med < best) {
          best = med;
          bestc = chunk_of[(size_t)k];
        }
      }
      if (have_dflt && best > 0.99 * dflt) {
        best = dflt;
        bestc = ch0;
      }
      tuned.push_back(TunedPass{mode, tr, out_rows, red_len, rb, bestc});
      int sb;
      int64_t cb;
      plan_pass_chunk(out_rows, red_len, rb, bestc, &sb, &cb);
      report[4 * dir + 0] = (double)cb;
      report[4 * dir + 1] = (double)sb;
      report[4 * dir + 2] = best;
      report[4 * dir + 3] = dflt;
    }
  }

  // rows kernel reading split partials of the current pass mode
  template <class F>
  void with_parts(void* part, F&& f) {
    if (pass_mode() == PASS_F64) {
      RowArgs<double> A;
      A.parts = (const double*)part;
      f(A);
    } else {
      RowArgs<float> A;
      A.parts = (const float*)part;
      f(A);
    }
  }

  // CholeskyQR2 of X (rows x r, ld rows) given the first kernel already
  // produced X and EPI_CHOL1 (R1 in small).  Writes Q to `out` (ld rows),
  // Rtot in small; optional packed output and cross reduction with W.
  //
  // passes = 3 is shifted CholeskyQR3: the first factor (EPI_CHOL1 with
  // shift_rows) came from G + s I, so Q1 = X R1^{-1} is only moderately
  // conditioned; two more CholeskyQR passes make it orthogonal to O(eps) for
  // cond(X) up to ~1/eps.  X is overwritten by the middle pass.
  //
  // With passes = 2 and an EPI_STILDE epilogue, the second Gram matrix and the
  // cross product Q1'W are reduced together and S~ is formed in the second
  // Cholesky epilogue (U1'W = R2^{-T} Q1'W), so the last row kernel needs no
  // reduction.
  //
  // first: the caller launched the first kernel with defer_epi and these
  // EPI_CHOL1 arguments (only where fold_chol(rr, passes)); then each Cholesky
  // epilogue runs in the prologue of the row kernel that applies its R^{-1},
  // which saves one launch per factor.
  static bool fold_chol(int rr, int passes = 2) { return passes == 2 && rr <= 32; }
  void cholqr_tail(double* X, double* tmp, int64_t rows_, int rr, double* out, bool pk_forN, const double* W,
                   int rw, EpiArgs final_epi, int sfinal, int passes = 2, const EpiArgs* first = nullptr) {
    const bool fuse_st = passes == 2 && W && final_epi.kind == EPI_STILDE && rw == rr && rr <= 32;
    REQUIRE(!first || fold_chol(rr, passes), "internal: deferred Cholesky epilogue cannot be folded");
    RowArgs<double> B;
    B.rows = rows_;
    B.r_out = rr;
    B.r_in = rr;
    B.X = X;
    B.ldx = rows_;
    B.Y = small->R1inv;
    B.ldy = rr;
    B.out = tmp;
    B.ldo = rows_;
    B.reduce = 1;
    B.epi.kind = EPI_CHOL2;
    B.epi.breakdown_is_error = final_epi.breakdown_is_error;
    B.epi.S = (sfinal && passes == 2) ? S : nullptr;
    B.epi.ld_s = (int)rcap;
    B.epi.transpose = sfinal == 2;
    if (fuse_st) {
      B.W = W;
      B.ldw = rows_;
      B.rw = rw;
      B.gram_w = 1;
      B.epi.stilde = final_epi.transpose ? 2 : 1;
    }
    if (first) {
      B.pre = *first;
      B.pre_C = small->Cred;
      B.pre_a = rr;
      B.pre_b = rr;
      B.defer_epi = 1;
    }
    rows(B);
    const double* last = tmp;
    if (passes == 3) {
      RowArgs<double> B3;  // X = tmp R2^{-1}; its Gram -> R3, Rtot = R3 R2 R1
      B3.rows = rows_;
      B3.r_out = rr;
      B3.r_in = rr;
      B3.X = tmp;
      B3.ldx = rows_;
      B3.Y = small->R2inv;
      B3.ldy = rr;
      B3.out = X;
      B3.ldo = rows_;
      B3.reduce = 1;
      B3.epi.kind = EPI_CHOL3;
      B3.epi.breakdown_is_error = final_epi.breakdown_is_error;
      B3.epi.S = sfinal ? S : nullptr;
      B3.epi.ld_s = (int)rcap;
      B3.epi.transpose = sfinal == 2;
      rows(B3);
      last = X;
    }
    RowArgs<double> C;
    C.rows = rows_;
    C.r_out = rr;
    C.r_in = rr;
    C.X = last;
    C.ldx = rows_;
    C.Y = small->R2inv;
    C.ldy = rr;
    C.out = out;
    C.ldo = rows_;
    if (first) {
      C.pre = B.epi;
      C.pre_C = small->Cred;
      C.pre_a = rr;
      C.pre_b = fuse_st ? 2 * rr : rr;
    }
    set_pack_outputs(C, pk_forN, pass_mode(), rb_for(pass_mode(), rr));
    if (!fuse_st && (final_epi.kind != EPI_NONE || W)) {
      C.reduce = 1;
      C.W = W;
      C.ldw = rows_;
      C.rw = rw;
      C.epi = final_epi;
    }
    rows(C);
  }

  // After a KSL CholeskyQR2 of X (rows_ x rr) into Q: when a pivot was dropped
  // (a rank-deficient K or L), the QR is recomputed from X by reorthogonalised
  // Gram-Schmidt with an orthonormal completion, as Householder QR (qr! in
  // projector_splitting.jl:137,149,174,186) returns it, and the products of its
  // R are rewritten: M = R - Q' W (mt: transposed) and Z = R (zt: R'); the
  // packed copy of Q for the next pass is updated.  One one-workgroup launch
  // that returns at once when nothing was dropped.
  void complete_dropped(const double* X, double* Q, int64_t rows_, int rr, bool pk_forN, double* M, int ldm, int mt,
                        const double* W, int64_t ldw, double* Z, int ldz, int zt) {
    CompleteArgs C;
    C.X = X;
    C.Q = Q;
    C.rows = rows_;
    C.k = rr;
    C.Rinv1 = small->R1inv;
    C.Rinv2 = small->R2inv;
    if (pack_enabled) {
      const int mode = pass_mode();
      C.pk_cols = rb_for(mode, rr);
      if (mode == PASS_F64) {
        C.pk_f64 = pk_forN ? pkN_f64 : pkT_f64;
      } else if (mode == PASS_F32) {
        C.pk_f32 = pk_forN ? pkN_f32 : pkT_f32;
        C.pk_layout = (pk_forN != tview) ? 0 : 1;
      } else {
        C.pk_hi = pk_forN ? pkN_hi : pkT_hi;
        C.pk_lo = pk_forN ? pkN_lo : pkT_lo;
        C.pk_layout = (pk_forN != tview) ? 0 : 1;
      }
    }
    C.M = M;
    C.ldm = ldm;
    C.mt = mt;
    C.W = W;
    C.ldw = ldw;
    C.Z = Z;
    C.ldz = ldz;
    C.zt = zt;
    timed(LRI_KERNEL_SMALL, [&] { return launch_complete_columns(C, stream); });
  }

  void check_step(int64_t kp, int64_t kn) {
    REQUIRE(r >= 1, "state not set (call lri_set_state)");
    REQUIRE(!snaps.empty(), "no snapshots (call lri_snapshots_reserve/upload or lri_snapshots_bind)");
    REQUIRE(kp >= 0 && kp < (int64_t)snaps.size() && kn >= 0 && kn < (int64_t)snaps.size(),
            "snapshot index out of range");
    REQUIRE(snaps[kp] && snaps[kn], "snapshot slot is empty");
  }

  // ------------------------------------------------------------------ steps
  // primal_LT_step! (projector_splitting.jl:129-152)
  void ksl_primal(int64_t kp, int64_t kn) {
    const int rr = (int)r;
    const bool fold = fold_chol(rr);
    EpiArgs c1;
    c1.kind = EPI_CHOL1;
    c1.breakdown_is_error = 0;  // dropped columns are completed
    ensure_packed(true);
    const int sN = passN(kp, kn, rr, partN);
    with_parts(partN, [&](auto A) {  // K = U S + dA V ; P = dA V ; Gram -> R1
      A.rows = n;
      A.nsplit = sN;
      A.rb = rb_for(pass_mode(), rr);
      A.r_out = rr;
      A.r_in = rr;
      A.sum_out = P;
      A.ld_sum = n;
      A.X = U;
      A.ldx = n;
      A.Y = S;
      A.ldy = (int)rcap;
      A.out = K;
      A.ldo = n;
      A.reduce = 1;
      A.epi = c1;
      A.defer_epi = fold;
      rows(A);
    });
    EpiArgs e;
    e.kind = EPI_STILDE;  // S~ = R - U1' P
    e.breakdown_is_error = 0;
    cholqr_tail(K, K2, n, rr, U, false, P, rr, e, 0, 2, fold ? &c1 : nullptr);
    complete_dropped(K, U, n, rr, false, small->St, rr, 0, P, n, nullptr, 0, 0);  // S~ = R - U1' P
    const int sT = passT(kp, kn, rr, partT);
    with_parts(partT, [&](auto A) {  // L = V S~' + dA' U1 ; Gram -> R1
      A.rows = m;
      A.nsplit = sT;
      A.rb = rb_for(pass_mode(), rr);
      A.r_out = rr;
      A.r_in = rr;
      A.X = V;
      A.ldx = m;
      A.Y = small->St;
      A.ldy = rr;
      A.transY = 1;
      A.out = L;
      A.ldo = m;
      A.reduce = 1;
      A.epi = c1;
      A.defer_epi = fold;
      rows(A);
    });
    EpiArgs none;
    none.breakdown_is_error = 0;
    cholqr_tail(L, L2, m, rr, V, true, nullptr, 0, none, 2, 2, fold ? &c1 : nullptr);  // V1, S = R'
    complete_dropped(L, V, m, rr, true, nullptr, 0, 0, nullptr, 0, S, (int)rcap, 1);  // S = R'
  }

  // dual_LT_step! (projector_splitting.jl:166-189)
  void ksl_dual(int64_t kp, int64_t kn) {
    const int rr = (int)r;
    const bool fold = fold_chol(rr);
    EpiArgs c1;
    c1.kind = EPI_CHOL1;
    c1.breakdown_is_error = 0;  // dropped columns are completed
    ensure_packed(false);
    const int sT = passT(kp, kn, rr, partT);
    with_parts(partT, [&](auto A) {  // L = V S' + dA' U ; Q = dA' U
      A.rows = m;
      A.nsplit = sT;
      A.rb = rb_for(pass_mode(), rr);
      A.r_out = rr;
      A.r_in = rr;
      A.sum_out = Q;
      A.ld_sum = m;
      A.X = V;
      A.ldx = m;
      A.Y = S;
      A.ldy = (int)rcap;
      A.transY = 1;
      A.out = L;
      A.ldo = m;
      A.reduce = 1;
      A.epi = c1;
      A.defer_epi = fold;
      rows(A);
    });
    EpiArgs e;
    e.kind = EPI_STILDE;  // S~ = (R - V1' Q)' = R' - U' dA V1
    e.transpose = 1;
    e.breakdown_is_error = 0;
    cholqr_tail(L, L2, m, rr, V, true, Q, rr, e, 0, 2, fold ? &c1 : nullptr);
    complete_dropped(L, V, m, rr, true, small->St, rr, 1, Q, m, nullptr, 0, 0);  // S~ = (R - V1' Q)'
    const int sN = passN(kp, kn, rr, partN);
    with_parts(partN, [&](auto A) {  // K = U S~ + dA V1
      A.rows = n;
      A.nsplit = sN;
      A.rb = rb_for(pass_mode(), rr);
      A.r_out = rr;
      A.r_in = rr;
      A.X = U;
      A.ldx = n;
      A.Y = small->St;
      A.ldy = rr;
      A.out = K;
      A.ldo = n;
      A.reduce = 1;
      A.epi = c1;
      A.defer_epi = fold;
      rows(A);
    });
    EpiArgs none;
    none.breakdown_is_error = 0;
    cholqr_tail(K, K2, n, rr, U, false, nullptr, 0, none, 1, 2, fold ? &c1 : nullptr);  // U1, S = R
    complete_dropped(K, U, n, rr, false, nullptr, 0, 0, nullptr, 0, S, (int)rcap, 0);  // S = R
  }

  // unconventional_step! (unconventional.jl:133-157)
  void bug(int64_t kp, int64_t kn) {
    const int rr = (int)r;
    const bool fold = fold_chol(rr);
    EpiArgs c1;
    c1.kind = EPI_CHOL1;
    ensure_packed(true);
    ensure_packed(false);
    const int sN = passN(kp, kn, rr, partN);  // dA V0
    const int sT = passT(kp, kn, rr, partT);  // dA' U0
    with_parts(partN, [&](auto A) {
      A.rows = n;
      A.nsplit = sN;
      A.rb = rb_for(pass_mode(), rr);
      A.r_out = rr;
      A.r_in = rr;
      A.X = U;
      A.ldx = n;
      A.Y = S;
      A.ldy = (int)rcap;
      A.out = K;
      A.ldo = n;
      A.reduce = 1;
      A.epi = c1;
      A.defer_epi = fold;
      rows(A);
    });
    EpiArgs eM;
    eM.kind = EPI_BUG_M;  // M = U1' U0
    cholqr_tail(K, K2, n, rr, U2, false, U, rr, eM, 0, 2, fold ? &c1 : nullptr);
    with_parts(partT, [&](auto A) {
      A.rows = m;
      A.nsplit = sT;
      A.rb = rb_for(pass_mode(), rr);
      A.r_out = rr;
      A.r_in = rr;
      A.X = V;
      A.ldx = m;
      A.Y = S;
      A.ldy = (int)rcap;
      A.transY = 1;
      A.out = L;
      A.ldo = m;
      A.reduce = 1;
      A.epi = c1;
      A.defer_epi = fold;
      rows(A);
    });
    EpiArgs eN;
    eN.kind = EPI_BUG_N;  // N = V1' V0, Sinit = M S0 N'
    eN.S = S;
    eN.ld_s = (int)rcap;
    cholqr_tail(L, L2, m, rr, V2, true, V, rr, eN, 0, 2, fold ? &c1 : nullptr);
    const int sN2 = passN(kp, kn, rr, partN);  // dA V1
    with_parts(partN, [&](auto A) {  // S1 = Sinit + U1' dA V1
      A.rows = n;
      A.nsplit = sN2;
      A.rb = rb_for(pass_mode(), rr);
      A.r_out = rr;
      A.reduce = 1;
      A.W = U2;
      A.ldw = n;
      A.rw = rr;
      A.epi.kind = EPI_BUG_S;
      A.epi.S = S;
      A.epi.ld_s = (int)rcap;
      rows(A);
    });
    std::swap(U, U2);
    std::swap(V, V2);
  }

  // greedy_step! on an SVD-like state (greedy_integrator.jl:94-104) with
  // X = A[kn] (update_data!, data_integrator.jl:22-28): U1 = orth(X V0),
  // V1 = orth(X' U0), S1 = U1' X V1.  The QRs are shifted CholeskyQR3 (the
  // products carry the snapshot's whole spectrum); dependent directions become
  // zero columns.
  void greedy(int64_t kn) {
    const int rr = (int)r;
    const int rb = rb_for(pass_mode(), rr);
    const EpiArgs none = de_epi();
    ensure_packed(true);
    ensure_packed(false);
    const int sN = passN(-1, kn, rr, partN);  // X V0
    const int sT = passT(-1, kn, rr, partT);  // X' U0
    with_parts(partN, [&](auto A) {
      A.rows = n;
      A.nsplit = sN;
      A.rb = rb;
      A.r_out = rr;
      A.out = K;
      A.ldo = n;
      A.reduce = 1;
      A.epi = shifted_chol(n);
      rows(A);
    });
    cholqr_tail(K, K2, n, rr, U2, false, nullptr, 0, none, 0, 3);  // U1 (packed for pass T)
    with_parts(partT, [&](auto A) {
      A.rows = m;
      A.nsplit = sT;
      A.rb = rb;
      A.r_out = rr;
      A.out = L;
      A.ldo = m;
      A.reduce = 1;
      A.epi = shifted_chol(m);
      rows(A);
    });
    cholqr_tail(L, L2, m, rr, V2, true, nullptr, 0, none, 0, 3);  // V1 (packed for pass N)
    const int sN2 = passN(-1, kn, rr, partN);  // X V1
    with_parts(partN, [&](auto A) {  // C = (X V1)' U1 = S1'
      A.rows = n;
      A.nsplit = sN2;
      A.rb = rb;
      A.r_out = rr;
      A.reduce = 1;
      A.W = U2;
      A.ldw = n;
      A.rw = rr;
      A.epi.kind = EPI_COPY;
      A.epi.transpose = 1;
      A.epi.dst = S;
      A.epi.ld_dst = (int)rcap;
      rows(A);
    });
    std::swap(U, U2);
    std::swap(V, V2);
  }

  // greedy_step! on a two-factor state U Z' (greedy_integrator.jl:84-92), Z in
  // the V buffer: Z1 = X' U0, then U1 = Q P' from the SVD X Z1 = Q Sigma P',
  // the orthogonal polar factor of X Z1, formed as X Z1 = Qc R (shifted
  // CholeskyQR3), R = Ur Sigma P' (Jacobi), U1 = Qc (Ur P').
  void greedy_twofactor(int64_t kn) {
    const int rr = (int)r;
    const int rb = rb_for(pass_mode(), rr);
    ensure_packed(false);
    const int sT = passT(-1, kn, rr, partT);  // Z1 = X' U0
    with_parts(partT, [&](auto A) {
      A.rows = m;
      A.nsplit = sT;
      A.rb = rb;
      A.r_out = rr;
      A.out = V;
      A.ldo = m;
      set_pack_outputs(A, true, pass_mode(), rb);
      rows(A);
    });
    greedy_polar(kn);
  }

  // greedy_step! on a MatrixHybridProblem (greedy_integrator.jl:73-82), the
  // two-factor state U Z' with Z in the V slot: the Z integrator advances Z over
  // dt with U held, Z' = f(Z, U, t) = F(U Z')' U -- for the operator descriptor
  // the L sub-problem of the matrix-ODE steps, Z' = B Z + Z (U'AU) + H (G'U)
  // (- Q(Z) with the tensor W = (D U) x U x U for the quadratic term), N RK4
  // steps -- then X = A[kn] and U1 = the polar factor of X Z1 (greedy_polar).
  void hybrid_greedy(int64_t kn, double dt, int64_t nsub) {
    const int rr = (int)r;
    const double h = dt / (double)nsub;
    hipLaunchKernelGGL(k_eye, dim3(1), dim3(256), 0, stream, de->Snew, rr);
    HIPCHECK(hipGetLastError());
    pack_enabled = false;
    if (op_quad) {
      quad_galerkin(U, true, rr);                      // U'AU, G'U, W(U)
      quad_init(V, m, rr, de->Snew, rr, 1, L);         // L = Z
      quad_rows(false, rr, L, L2, Q, h, nsub);
      HIPCHECK(hipMemcpyAsync(V, L, sizeof(double) * m * rr, hipMemcpyDeviceToDevice, stream));
    } else {
      de_galerkin(U, true, rr, de->Ga, de->gaT);       // U'AU, G'U
      de_eig(de->Ga, rr, de->Wa, de->alpha);
      de_substep(V, false, rr, de->Snew, rr, 1, de->Wa, de->alpha, de->gaT, h, nsub, V, L2, de_epi());
    }
    pack_enabled = true;
    pkN_mode = pkT_mode = -1;
    pack(V, m, m, rr, true);
    greedy_polar(kn);
  }

  // U1 = Q P' for the SVD X Z1 = Q Sigma P' (greedy_integrator.jl:79-81,89-91), Z1
  // in V and packed for pass N: X Z1 = Qc R (shifted CholeskyQR3), R = Ur Sigma P'
  // (Jacobi), U1 = Qc (Ur P').
  void greedy_polar(int64_t kn) {
    const int rr = (int)r;
    const int rb = rb_for(pass_mode(), rr);
    const EpiArgs none = de_epi();
    const int sN = passN(-1, kn, rr, partN);  // X Z1
    with_parts(partN, [&](auto A) {
      A.rows = n;
      A.nsplit = sN;
      A.rb = rb;
      A.r_out = rr;
      A.out = K;
      A.ldo = n;
      A.reduce = 1;
      A.epi = shifted_chol(n);
      rows(A);
    });
    cholqr_tail(K, K2, n, rr, U2, false, nullptr, 0, none, 0, 3);  // Qc, R = Rtot
    SvdArgs sv;
    sv.S = small->Rtot;
    sv.lds = rr;
    sv.k = rr;
    sv.U = small->Ush;
    sv.V = small->Vsh;
    sv.sigma = small->sigma;
    sv.r_out = &small->r_new;
    sv.tol = 0.0;
    sv.rmax = rr;
    sv.work = svd_work;
    timed(LRI_KERNEL_SMALL, [&] { return launch_jacobi_svd(sv, stream); });
    timed(LRI_KERNEL_SMALL, [&] { return launch_small_abt(small->Ush, small->Vsh, rr, small->Ctmp, stream); });
    RowArgs<double> B;  // U1 = Qc (Ur P')
    B.rows = n;
    B.r_out = rr;
    B.r_in = rr;
    B.X = U2;
    B.ldx = n;
    B.Y = small->Ctmp;
    B.ldy = rr;
    B.out = U;
    B.ldo = n;
    set_pack_outputs(B, false, pass_mode(), rb);
    rows(B);
  }

  // ------------------------------------------------------------- adjoint
  // Reverse of unconventional_step! on data (unconventional.jl:133-157): with the
  // state (U0, S0, V0) of the step and the cotangents (Ub, Sb, Vb) of its result
  // (U1, S1, V1) in P[:, :r], adj->Sb1 and Q[:, :r], leaves the cotangents of
  // (U0, S0, V0) in P[:, :r], adj->Sb0 and Q[:, :r].  The forward step is
  // recomputed up to U1, V1, the R factors, M and N (S1 is not needed); then
  //   Mb = Sb1 N S0', Nb = Sb1' M S0, Sb0 = M' Sb1 N,
  //   Ub1 += dA V1 Sb1' + U0 Mb',  Vb1 += dA' U1 Sb1 + V0 Nb',
  //   Kbar = (Ub1 + U1 copyltu(-Ub1'U1)) RK^{-T},  Lbar likewise,
  //   Ub0 = U1 Mb + Kbar S0' + dA Lbar,  Vb0 = V1 Nb + Lbar S0 + dA' Kbar,
  //   Sb0 += U0' Kbar + Lbar' V0.
  // Four passes over dA (two forward, two reverse).
  void bug_adjoint(int64_t kp, int64_t kn) {
    const int rr = (int)r;
    const int mode = pass_mode();
    const int rb = rb_for(mode, rr);
    const size_t rsz = sizeof(double) * rr * rr;
    // forward: K1 = U0 S0 + dA V0 -> U1, RK ; L1 = V0 S0' + dA' U0 -> V1, RL ; M, N
    ensure_packed(true);
    ensure_packed(false);
    const int sN = passN(kp, kn, rr, partN);
    const int sT = passT(kp, kn, rr, partT);
    with_parts(partN, [&](auto A) {
      A.rows = n;
      A.nsplit = sN;
      A.rb = rb;
      A.r_out = rr;
      A.r_in = rr;
      A.X = U;
      A.ldx = n;
      A.Y = S;
      A.ldy = (int)rcap;
      A.out = K;
      A.ldo = n;
      A.reduce = 1;
      A.epi.kind = EPI_CHOL1;
      rows(A);
    });
    EpiArgs eM;
    eM.kind = EPI_BUG_M;  // M = U1' U0
    cholqr_tail(K, K2, n, rr, U2, false, U, rr, eM, 0);
    HIPCHECK(hipMemcpyAsync(adj->RK, small->Rtot, rsz, hipMemcpyDeviceToDevice, stream));
    with_parts(partT, [&](auto A) {
      A.rows = m;
      A.nsplit = sT;
      A.rb = rb;
      A.r_out = rr;
      A.r_in = rr;
      A.X = V;
      A.ldx = m;
      A.Y = S;
      A.ldy = (int)rcap;
      A.transY = 1;
      A.out = L;
      A.ldo = m;
      A.reduce = 1;
      A.epi.kind = EPI_CHOL1;
      rows(A);
    });
    EpiArgs eNs;
    eNs.kind = EPI_COPY;  // N = V1' V0
    eNs.dst = small->N;
    cholqr_tail(L, L2, m, rr, V2, true, V, rr, eNs, 0);
    HIPCHECK(hipMemcpyAsync(adj->RL, small->Rtot, rsz, hipMemcpyDeviceToDevice, stream));
    timed(LRI_KERNEL_SMALL, [&] { return launch_bug_adj_pre(adj, small->M, small->N, S, (int)rcap, rr, stream); });
    double* Ub = P;           // in: Ub1, out: Ub0
    double* Ubf = P + n * rr;  // Ub1 + dA V1 Sb1' + U0 Mb'
    double* Vb = Q;
    double* Vbf = Q + m * rr;
    // pass operands: V1 Sb1' (N side), U1 Sb1 (T side)
    for (int side = 0; side < 2; ++side) {
      RowArgs<double> A;
      A.rows = side == 0 ? m : n;
      A.r_out = rr;
      A.r_in = rr;
      A.X = side == 0 ? V2 : U2;
      A.ldx = A.rows;
      A.Y = adj->Sb1;
      A.ldy = rr;
      A.transY = side == 0;
      set_pack_outputs(A, side == 0, mode, rb);
      rows(A);
    }
    const int sN1 = passN(kp, kn, rr, partN);  // dA V1 Sb1'
    const int sT1 = passT(kp, kn, rr, partT);  // dA' U1 Sb1
    with_parts(partN, [&](auto A) {  // Ubf = parts + Ub + U0 Mb' ; C = Ubf' U1 -> CqK
      A.rows = n;
      A.nsplit = sN1;
      A.rb = rb;
      A.r_out = rr;
      A.r_in = rr;
      A.base = Ub;
      A.ld_base = n;
      A.X = U;
      A.ldx = n;
      A.Y = adj->Mb;
      A.ldy = rr;
      A.transY = 1;
      A.out = Ubf;
      A.ldo = n;
      A.reduce = 1;
      A.W = U2;
      A.ldw = n;
      A.rw = rr;
      A.epi.kind = EPI_COPYLTU;
      A.epi.dst = adj->CqK;
      rows(A);
    });
    with_parts(partT, [&](auto A) {  // Vbf = parts + Vb + V0 Nb' ; C = Vbf' V1 -> CqL
      A.rows = m;
      A.nsplit = sT1;
      A.rb = rb;
      A.r_out = rr;
      A.r_in = rr;
      A.base = Vb;
      A.ld_base = m;
      A.X = V;
      A.ldx = m;
      A.Y = adj->Nb;
      A.ldy = rr;
      A.transY = 1;
      A.out = Vbf;
      A.ldo = m;
      A.reduce = 1;
      A.W = V2;
      A.ldw = m;
      A.rw = rr;
      A.epi.kind = EPI_COPYLTU;
      A.epi.dst = adj->CqL;
      rows(A);
    });
    // Kbar = (Ubf + U1 CqK) RK^{-T} (packed for pass T; C = Kbar' U0 -> CK), Lbar likewise
    for (int side = 0; side < 2; ++side) {
      const int64_t rows_ = side == 0 ? n : m;
      double* Z = side == 0 ? K2 : L2;
      double* bar = side == 0 ? K : L;
      RowArgs<double> A;
      A.rows = rows_;
      A.r_out = rr;
      A.r_in = rr;
      A.base = side == 0 ? Ubf : Vbf;
      A.ld_base = rows_;
      A.X = side == 0 ? U2 : V2;
      A.ldx = rows_;
      A.Y = side == 0 ? adj->CqK : adj->CqL;
      A.ldy = rr;
      A.out = Z;
      A.ldo = rows_;
      rows(A);
      RowArgs<double> B;
      B.rows = rows_;
      B.r_out = rr;
      B.r_in = rr;
      B.X = Z;
      B.ldx = rows_;
      B.Y = side == 0 ? adj->RKinvT : adj->RLinvT;
      B.ldy = rr;
      B.out = bar;
      B.ldo = rows_;
      set_pack_outputs(B, side == 1, mode, rb);  // Kbar for pass T, Lbar for pass N
      B.reduce = 1;
      B.W = side == 0 ? U : V;
      B.ldw = rows_;
      B.rw = rr;
      B.epi.kind = EPI_COPY;
      B.epi.dst = side == 0 ? adj->CK : adj->CL;
      rows(B);
    }
    const int sN2 = passN(kp, kn, rr, partN);  // dA Lbar
    const int sT2 = passT(kp, kn, rr, partT);  // dA' Kbar
    for (int side = 0; side < 2; ++side) {
      const int64_t rows_ = side == 0 ? n : m;
      double* T0 = (side == 0 ? K2 : L2) + rows_ * rr;
      with_parts(side == 0 ? partN : partT, [&](auto A) {  // T0 = parts + (U1 Mb | V1 Nb)
        A.rows = rows_;
        A.nsplit = side == 0 ? sN2 : sT2;
        A.rb = rb;
        A.r_out = rr;
        A.r_in = rr;
        A.X = side == 0 ? U2 : V2;
        A.ldx = rows_;
        A.Y = side == 0 ? adj->Mb : adj->Nb;
        A.ldy = rr;
        A.out = T0;
        A.ldo = rows_;
        rows(A);
      });
      RowArgs<double> B;  // (Ub0 | Vb0) = T0 + (Kbar S0' | Lbar S0)
      B.rows = rows_;
      B.r_out = rr;
      B.r_in = rr;
      B.base = T0;
      B.ld_base = rows_;
      B.X = side == 0 ? K : L;
      B.ldx = rows_;
      B.Y = S;
      B.ldy = (int)rcap;
      B.transY = side == 0;
      B.out = side == 0 ? Ub : Vb;
      B.ldo = rows_;
      rows(B);
    }
    timed(LRI_KERNEL_SMALL, [&] { return launch_bug_adj_post(adj, rr, stream); });
    pkN_mode = pkT_mode = -1;  // the panels hold adjoint operands now
  }

  // Reverse of primal_LT_step! on data (projector_splitting.jl:129-152), same
  // conventions as bug_adjoint.  Forward recompute: P = dA V0, K = U0 S0 + P ->
  // (U1, R1), S~ = R1 - U1'P, L = V0 S~' + dA'U1 -> (V1, R2); S1 = R2'.  Reverse
  // (both QR factors carry cotangents: Kbar = (Qbar + Q copyltu(R Rbar' - Qbar'Q)) R^{-T}):
  //   Lbar from (Vb1, R2bar = Sb1');  S~bar = Lbar' V0;  Vb0 = Lbar S~ + dA'(Kbar - U1 S~bar);
  //   Ub1 += dA Lbar - P S~bar';  Kbar from (Ub1, R1bar = S~bar);  Ub0 = Kbar S0';  Sb0 = U0'Kbar.
  void ksl_adjoint(int64_t kp, int64_t kn) {
    const int rr = (int)r;
    const int mode = pass_mode();
    const int rb = rb_for(mode, rr);
    const size_t rsz = sizeof(double) * rr * rr;
    ensure_packed(true);
    const int sN = passN(kp, kn, rr, partN);
    with_parts(partN, [&](auto A) {  // K = U S + dA V ; P = dA V ; Gram -> R1
      A.rows = n;
      A.nsplit = sN;
      A.rb = rb;
      A.r_out = rr;
      A.r_in = rr;
      A.sum_out = P;
      A.ld_sum = n;
      A.X = U;
      A.ldx = n;
      A.Y = S;
      A.ldy = (int)rcap;
      A.out = K;
      A.ldo = n;
      A.reduce = 1;
      A.epi.kind = EPI_CHOL1;
      rows(A);
    });
    EpiArgs e;
    e.kind = EPI_STILDE;  // S~ = R - U1' P
    cholqr_tail(K, K2, n, rr, U2, false, P, rr, e, 0);  // U1 (packed for pass T)
    HIPCHECK(hipMemcpyAsync(adj->RK, small->Rtot, rsz, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(adj->St, small->St, rsz, hipMemcpyDeviceToDevice, stream));
    const int sT = passT(kp, kn, rr, partT);
    with_parts(partT, [&](auto A) {  // L = V S~' + dA' U1 ; Gram -> R1
      A.rows = m;
      A.nsplit = sT;
      A.rb = rb;
      A.r_out = rr;
      A.r_in = rr;
      A.X = V;
      A.ldx = m;
      A.Y = small->St;
      A.ldy = rr;
      A.transY = 1;
      A.out = L;
      A.ldo = m;
      A.reduce = 1;
      A.epi.kind = EPI_CHOL1;
      rows(A);
    });
    EpiArgs none;
    cholqr_tail(L, L2, m, rr, V2, true, nullptr, 0, none, 0);  // V1
    HIPCHECK(hipMemcpyAsync(adj->RL, small->Rtot, rsz, hipMemcpyDeviceToDevice, stream));
    timed(LRI_KERNEL_SMALL, [&] { return launch_ksl_adj_small(adj, rr, 0, stream); });
    double* Ub = P + n * rr;   // P[:, :r] holds dA V0; the cotangents sit in the next r columns
    double* Vb = Q;
    double* Vbf = Q + m * rr;
    (void)Vbf;
    // CqL = copyltu(R2 Sb1 - Vb1' V1)
    {
      RowArgs<double> A;
      A.rows = m;
      A.r_out = rr;
      A.base = Vb;
      A.ld_base = m;
      A.reduce = 1;
      A.W = V2;
      A.ldw = m;
      A.rw = rr;
      A.epi.kind = EPI_COPYLTU;
      A.epi.S = adj->Mb;
      A.epi.ld_s = rr;
      A.epi.dst = adj->CqL;
      rows(A);
    }
    // Lbar = (Vb1 + V1 CqL) R2^{-T}, packed for pass N; C = Lbar' V0 = S~bar -> CL
    {
      RowArgs<double> A;
      A.rows = m;
      A.r_out = rr;
      A.r_in = rr;
      A.base = Vb;
      A.ld_base = m;
      A.X = V2;
      A.ldx = m;
      A.Y = adj->CqL;
      A.ldy = rr;
      A.out = L2;
      A.ldo = m;
      rows(A);
      RowArgs<double> B;
      B.rows = m;
      B.r_out = rr;
      B.r_in = rr;
      B.X = L2;
      B.ldx = m;
      B.Y = adj->RLinvT;
      B.ldy = rr;
      B.out = L;
      B.ldo = m;
      set_pack_outputs(B, true, mode, rb);
      B.reduce = 1;
      B.W = V;
      B.ldw = m;
      B.rw = rr;
      B.epi.kind = EPI_COPY;
      B.epi.dst = adj->CL;
      rows(B);
    }
    timed(LRI_KERNEL_SMALL, [&] { return launch_ksl_adj_small(adj, rr, 1, stream); });  // Nb = R1 S~bar'
    const int sN1 = passN(kp, kn, rr, partN);  // dA Lbar
    with_parts(partN, [&](auto A) {  // Ubf = parts + Ub1 - P S~bar' ; CqK = copyltu(R1 S~bar' - Ubf' U1)
      A.rows = n;
      A.nsplit = sN1;
      A.rb = rb;
      A.r_out = rr;
      A.r_in = rr;
      A.base = Ub;
      A.ld_base = n;
      A.X = P;
      A.ldx = n;
      A.Y = adj->CL;
      A.ldy = rr;
      A.transY = 1;
      A.xy_sign = -1.0;
      A.out = K2;
      A.ldo = n;
      A.reduce = 1;
      A.W = U2;
      A.ldw = n;
      A.rw = rr;
      A.epi.kind = EPI_COPYLTU;
      A.epi.S = adj->Nb;
      A.epi.ld_s = rr;
      A.epi.dst = adj->CqK;
      rows(A);
    });
    // Kbar = (Ubf + U1 CqK) R1^{-T} -> K ; C = Kbar' U0 -> CK
    {
      RowArgs<double> A;
      A.rows = n;
      A.r_out = rr;
      A.r_in = rr;
      A.base = K2;
      A.ld_base = n;
      A.X = U2;
      A.ldx = n;
      A.Y = adj->CqK;
      A.ldy = rr;
      A.out = K2 + n * rr;
      A.ldo = n;
      rows(A);
      RowArgs<double> B;
      B.rows = n;
      B.r_out = rr;
      B.r_in = rr;
      B.X = K2 + n * rr;
      B.ldx = n;
      B.Y = adj->RKinvT;
      B.ldy = rr;
      B.out = K;
      B.ldo = n;
      B.reduce = 1;
      B.W = U;
      B.ldw = n;
      B.rw = rr;
      B.epi.kind = EPI_COPY;
      B.epi.dst = adj->CK;
      rows(B);
    }
    // pass T operand Kbar - U1 S~bar
    {
      RowArgs<double> A;
      A.rows = n;
      A.r_out = rr;
      A.r_in = rr;
      A.base = K;
      A.ld_base = n;
      A.X = U2;
      A.ldx = n;
      A.Y = adj->CL;
      A.ldy = rr;
      A.xy_sign = -1.0;
      set_pack_outputs(A, false, mode, rb);
      rows(A);
    }
    const int sT1 = passT(kp, kn, rr, partT);  // dA'(Kbar - U1 S~bar)
    with_parts(partT, [&](auto A) {  // Vb0 = parts + Lbar S~
      A.rows = m;
      A.nsplit = sT1;
      A.rb = rb;
      A.r_out = rr;
      A.r_in = rr;
      A.X = L;
      A.ldx = m;
      A.Y = adj->St;
      A.ldy = rr;
      A.out = Vb;
      A.ldo = m;
      rows(A);
    });
    {
      RowArgs<double> A;  // Ub0 = Kbar S0'
      A.rows = n;
      A.r_out = rr;
      A.r_in = rr;
      A.X = K;
      A.ldx = n;
      A.Y = S;
      A.ldy = (int)rcap;
      A.transY = 1;
      A.out = Ub;
      A.ldo = n;
      rows(A);
    }
    timed(LRI_KERNEL_SMALL, [&] { return launch_ksl_adj_small(adj, rr, 2, stream); });  // Sb0 = CK'
    pkN_mode = pkT_mode = -1;
  }

  // The BUG or primal KSL reverse step with host cotangents: those of (U1, S1,
  // V1) in, those of (U0, S0, V0) out (the state is unchanged).
  void data_adjoint(bool ksl, int64_t kp, int64_t kn, double* Ub, int64_t ldub, double* Sb, int64_t ldsb, double* Vb,
                    int64_t ldvb) {
    const size_t d = sizeof(double);
    double* ubuf = ksl ? P + n * r : P;
    HIPCHECK(hipMemcpy2DAsync(ubuf, d * n, Ub, d * ldub, d * n, r, hipMemcpyDefault, stream));
    HIPCHECK(hipMemcpy2DAsync(Q, d * m, Vb, d * ldvb, d * m, r, hipMemcpyDefault, stream));
    HIPCHECK(hipMemcpy2DAsync(adj->Sb1, d * r, Sb, d * ldsb, d * r, r, hipMemcpyDefault, stream));
    if (ksl)
      ksl_adjoint(kp, kn);
    else
      bug_adjoint(kp, kn);
    HIPCHECK(hipMemcpy2DAsync(Ub, d * ldub, ubuf, d * n, d * n, r, hipMemcpyDefault, stream));
    HIPCHECK(hipMemcpy2DAsync(Vb, d * ldvb, Q, d * m, d * m, r, hipMemcpyDefault, stream));
    HIPCHECK(hipMemcpy2DAsync(Sb, d * ldsb, adj->Sb0, d * r, d * r, r, hipMemcpyDefault, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }

  // dual_LT_step! (projector_splitting.jl:166-189) is primal_LT_step! of the
  // transposed problem: dual(U, S, V; dA) = (V1', S1'', U1') for
  // primal(V, S', U; dA') = (U1', S1', V1').  Its reverse step is the primal one
  // run in the transposed view, with the U and V cotangents exchanged and the S
  // cotangents transposed on the way in and out.
  void ksl_dual_adjoint(int64_t kp, int64_t kn, double* Ub, int64_t ldub, double* Sb, int64_t ldsb, double* Vb,
                        int64_t ldvb) {
    // Sb may be host or device memory: it is transposed through a host copy
    const int rr = (int)r;
    const size_t d = sizeof(double);
    std::vector<double> sb0((size_t)rr * rr), sbt((size_t)rr * rr);
    HIPCHECK(hipMemcpy2D(sb0.data(), d * rr, Sb, d * ldsb, d * rr, rr, hipMemcpyDefault));
    for (int j = 0; j < rr; ++j)
      for (int i = 0; i < rr; ++i) sbt[i + (size_t)j * rr] = sb0[j + (size_t)i * rr];
    transpose_S();
    swap_sides();
    try {
      data_adjoint(true, kp, kn, Vb, ldvb, sbt.data(), rr, Ub, ldub);
    } catch (...) {
      swap_sides();
      try {
        transpose_S();
      } catch (...) {
      }
      throw;
    }
    swap_sides();
    transpose_S();
    for (int j = 0; j < rr; ++j)
      for (int i = 0; i < rr; ++i) sb0[i + (size_t)j * rr] = sbt[j + (size_t)i * rr];
    HIPCHECK(hipMemcpy2D(Sb, d * ldsb, sb0.data(), d * rr, d * rr, rr, hipMemcpyDefault));
  }

  // a device copy of the state (U, S, V, rank), for reverse steps that recompute a forward half
  double* backup = nullptr;
  int64_t backup_r = 0;
  void save_state() {
    if (!backup) HIPCHECK(hipMalloc((void**)&backup, sizeof(double) * (size_t)traj_stride()));
    const size_t d = sizeof(double);
    HIPCHECK(hipMemcpyAsync(backup, U, d * n * rcap, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(backup + n * rcap, S, d * rcap * rcap, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(backup + n * rcap + rcap * rcap, V, d * m * rcap, hipMemcpyDeviceToDevice, stream));
    backup_r = r;
  }
  void restore_state() {
    const size_t d = sizeof(double);
    HIPCHECK(hipMemcpyAsync(U, backup, d * n * rcap, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(S, backup + n * rcap, d * rcap * rcap, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(V, backup + n * rcap + rcap * rcap, d * m * rcap, hipMemcpyDeviceToDevice, stream));
    r = backup_r;
    pkN_mode = pkT_mode = -1;
  }

  // Orthonormal completion of span(X) against an orthonormal B: Q_perp with
  // [B | Q_perp] orthonormal (dependent directions become zero columns).
  // X = first kernel's output (rows x rr), block Gram-Schmidt twice, then CholeskyQR2.
  void complete_basis(double* X, const double* Bm, int64_t rows_, int rr, double* t1, double* t2,
                      double* qout, EpiArgs fin) {
    // pass 1: C = X' B
    {
      RowArgs<double> A;
      A.rows = rows_;
      A.r_out = rr;
      A.base = X;
      A.ld_base = rows_;
      A.reduce = 1;
      A.W = Bm;
      A.ldw = rows_;
      A.rw = rr;
      A.epi.kind = EPI_STORE;  // C -> small->Ctmp
      rows(A);
    }
    for (int pass = 0; pass < 2; ++pass) {
      // t = src - B (B' src) = src - B * C'   (C = src' B)
      RowArgs<double> A;
      A.rows = rows_;
      A.r_out = rr;
      A.r_in = rr;
      A.base = pass == 0 ? X : t1;
      A.ld_base = rows_;
      A.X = Bm;
      A.ldx = rows_;
      A.Y = small->Ctmp;
      A.ldy = rr;
      A.transY = 1;
      A.xy_sign = -1.0;
      A.out = pass == 0 ? t1 : t2;
      A.ldo = rows_;
      A.reduce = 1;
      if (pass == 0) {
        A.W = Bm;
        A.ldw = rows_;
        A.rw = rr;
        A.epi.kind = EPI_STORE;
      } else {
        A.epi.kind = EPI_CHOL1;
        A.epi.breakdown_is_error = 0;
      }
      rows(A);
    }
    // CholeskyQR2 tail on t2 -> qout, cross with B -> fin (EPI_RA_M / EPI_RA_N)
    fin.breakdown_is_error = 0;
    RowArgs<double> B2;
    B2.rows = rows_;
    B2.r_out = rr;
    B2.r_in = rr;
    B2.X = t2;
    B2.ldx = rows_;
    B2.Y = small->R1inv;
    B2.ldy = rr;
    B2.out = t1;
    B2.ldo = rows_;
    B2.reduce = 1;
    B2.epi.kind = EPI_CHOL2;
    B2.epi.breakdown_is_error = 0;
    rows(B2);
    RowArgs<double> C;
    C.rows = rows_;
    C.r_out = rr;
    C.r_in = rr;
    C.X = t1;
    C.ldx = rows_;
    C.Y = small->R2inv;
    C.ldy = rr;
    C.out = qout;
    C.ldo = rows_;
    C.reduce = 1;
    C.W = Bm;
    C.ldw = rows_;
    C.rw = rr;
    C.epi = fin;
    rows(C);
  }

  // rankadaptive_unconventional_step! (rank_adaptive_unconventional.jl:194-233)
  int64_t bug_ra(int64_t kp, int64_t kn, double tol, int64_t rmax) {
    const int rr = (int)r;
    const int k2 = 2 * rr;
    ensure_packed(true);
    ensure_packed(false);
    const int sN = passN(kp, kn, rr, partN);  // dA V0
    const int sT = passT(kp, kn, rr, partT);  // dA' U0
    // K = U S + dA V  -> K ; Uhat = [U | Q_perp(K)]
    with_parts(partN, [&](auto A) {
      A.rows = n;
      A.nsplit = sN;
      A.rb = rb_for(pass_mode(), rr);
      A.r_out = rr;
      A.r_in = rr;
      A.X = U;
      A.ldx = n;
      A.Y = S;
      A.ldy = (int)rcap;
      A.out = K;
      A.ldo = n;
      rows(A);
    });
    HIPCHECK(hipMemcpyAsync(Uhat, U, sizeof(double) * n * rr, hipMemcpyDeviceToDevice, stream));
    EpiArgs eM;
    eM.kind = EPI_RA_M;
    complete_basis(K, U, n, rr, K2, K2 + n * rcap, Uhat + n * rr, eM);
    with_parts(partT, [&](auto A) {
      A.rows = m;
      A.nsplit = sT;
      A.rb = rb_for(pass_mode(), rr);
      A.r_out = rr;
      A.r_in = rr;
      A.X = V;
      A.ldx = m;
      A.Y = S;
      A.ldy = (int)rcap;
      A.transY = 1;
      A.out = L;
      A.ldo = m;
      rows(A);
    });
    HIPCHECK(hipMemcpyAsync(Vhat, V, sizeof(double) * m * rr, hipMemcpyDeviceToDevice, stream));
    EpiArgs eN;
    eN.kind = EPI_RA_N;
    eN.S = S;
    eN.ld_s = (int)rcap;
    complete_basis(L, V, m, rr, L2, L2 + m * rcap, Vhat + m * rr, eN);
    // Shat = M S N' + Uhat' dA Vhat
    pack(Vhat, m, m, k2, true);
    const int sN2 = passN(kp, kn, k2, partN);
    with_parts(partN, [&](auto A) {
      A.rows = n;
      A.nsplit = sN2;
      A.rb = rb_for(pass_mode(), k2);
      A.r_out = k2;
      A.reduce = 1;
      A.W = Uhat;
      A.ldw = n;
      A.rw = k2;
      A.epi.kind = EPI_RA_S;
      rows(A);
    });
    return ra_finish(k2, tol, rmax);
  }

  // Reverse of rankadaptive_unconventional_step! on data
  // (rank_adaptive_unconventional.jl:194-233), the new rank r1 held fixed: with
  // the state (U0, S0, V0) of the step and the cotangents (Ub1, Sb1, Vb1) of its
  // result (n x r1, r1 x r1, m x r1) in P, radj->Sb1 and Q, leaves those of
  // (U0, S0, V0) in P[:, :r], radj->Sb0 and Q[:, :r].  The forward step is
  // recomputed up to the SVD Shat = Us Sig Vs' (three passes) with
  //   Uhat = [U0 | Q_U], Q_U R_U = Z_U = (I - U0 U0') K1,  K1 = U0 S0 + dA V0
  // (the reference's Q([K1 U0]) spans the same space, and U1 S1 V1' does not
  // depend on the basis); then
  //   Uhbar = Ub1 Us[:, :r1]',  Shbar = the SVD rule on (Uhat'Ub1, diag Sb1, Vhat'Vb1),
  //   Uhbar += U0 (S0 N' Shbar') + dA (Vhat Shbar'),  Vhbar likewise (two passes),
  //   Zbar = (Qbar + Q_U copyltu(-Qbar'Q_U)) R_U^{-T} for Qbar = Uhbar[:, r:],
  //   Kbar1 = Zbar - U0 (U0'Zbar),
  //   Ub0 = Uhbar[:, :r] + Uhat (Shbar N S0') - Zbar (K1'U0) - K1 (Zbar'U0) + Kbar1 S0' + dA Lbar1,
  //   Vb0 likewise, Sb0 = M' Shbar N (U0'Kbar1 and Lbar1'V0 vanish for orthonormal bases).
  // Seven passes over dA.  The rule needs distinct singular values of Shat and
  // full-rank completions.
  void bug_ra_adjoint(int64_t kp, int64_t kn, double tol, int64_t rmax, int r1) {
    const int rr = (int)r;
    const int k2 = 2 * rr;
    const int mode = pass_mode();
    const size_t rsz = sizeof(double) * rr * rr;
    RaAdjSmall* A = radj;
    // forward, as bug_ra up to the SVD (the state is left unchanged)
    ensure_packed(true);
    ensure_packed(false);
    const int sN = passN(kp, kn, rr, partN);
    const int sT = passT(kp, kn, rr, partT);
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      with_parts(ns ? partN : partT, [&](auto X) {  // K1 = U S + dA V ; L1 = V S' + dA' U
        X.rows = rows_;
        X.nsplit = ns ? sN : sT;
        X.rb = rb_for(mode, rr);
        X.r_out = rr;
        X.r_in = rr;
        X.X = ns ? U : V;
        X.ldx = rows_;
        X.Y = S;
        X.ldy = (int)rcap;
        X.transY = ns ? 0 : 1;
        X.out = ns ? K : L;
        X.ldo = rows_;
        rows(X);
      });
      HIPCHECK(hipMemcpyAsync(ns ? Uhat : Vhat, ns ? U : V, sizeof(double) * rows_ * rr, hipMemcpyDeviceToDevice,
                              stream));
      EpiArgs e;
      e.kind = ns ? EPI_RA_M : EPI_RA_N;
      if (!ns) {
        e.S = S;
        e.ld_s = (int)rcap;
      }
      double* t1 = ns ? K2 : L2;
      complete_basis(ns ? K : L, ns ? U : V, rows_, rr, t1, t1 + rows_ * rcap, (ns ? Uhat : Vhat) + rows_ * rr, e);
      HIPCHECK(hipMemcpyAsync(ns ? A->RU : A->RV, small->Rtot, rsz, hipMemcpyDeviceToDevice, stream));
    }
    pack(Vhat, m, m, k2, true);
    const int sN2 = passN(kp, kn, k2, partN);
    with_parts(partN, [&](auto X) {  // Shat = M S0 N' + Uhat' dA Vhat
      X.rows = n;
      X.nsplit = sN2;
      X.rb = rb_for(mode, k2);
      X.r_out = k2;
      X.reduce = 1;
      X.W = Uhat;
      X.ldw = n;
      X.rw = k2;
      X.epi.kind = EPI_RA_S;
      rows(X);
    });
    ra_svd(k2, tol, rmax);
    HIPCHECK(hipStreamSynchronize(stream));
    REQUIRE(*r_host == r1, "the cotangents' rank differs from the step's new rank (recomputed: " +
                               std::to_string(*r_host) + ")");
    // Uhbar = Ub1 Us[:, :r1]' (K2 / L2) and Uhat' Ub1 (radj->Ubs / Vbs)
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      RowArgs<double> X;
      X.rows = rows_;
      X.r_out = k2;
      X.r_in = r1;
      X.X = ns ? P : Q;
      X.ldx = rows_;
      X.Y = ns ? small->Ush : small->Vsh;
      X.ldy = k2;
      X.transY = 1;
      X.out = ns ? K2 : L2;
      X.ldo = rows_;
      rows(X);
      RowArgs<double> Y;
      Y.rows = rows_;
      Y.r_out = r1;
      Y.base = ns ? P : Q;
      Y.ld_base = rows_;
      Y.reduce = 1;
      Y.W = ns ? Uhat : Vhat;
      Y.ldw = rows_;
      Y.rw = k2;
      Y.epi.kind = EPI_COPY;
      Y.epi.transpose = 1;
      Y.epi.dst = ns ? A->Ubs : A->Vbs;
      rows(Y);
    }
    for (int stage = 0; stage < 5; ++stage)
      timed(LRI_KERNEL_SMALL, [&] { return launch_ra_adj_small(A, small, S, (int)rcap, rr, r1, stage, stream); });
    // pass operands Vhat Shbar' (N) and Uhat Shbar (T); Uhbar += parts + U0 G1 -> P, Vhbar likewise -> Q
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      RowArgs<double> X;
      X.rows = ns ? m : n;
      X.r_out = k2;
      X.r_in = k2;
      X.X = ns ? Vhat : Uhat;
      X.ldx = X.rows;
      X.Y = A->Shb;
      X.ldy = k2;
      X.transY = ns ? 1 : 0;
      set_pack_outputs(X, ns, mode, rb_for(mode, k2));
      rows(X);
    }
    const int sN3 = passN(kp, kn, k2, partN);
    const int sT3 = passT(kp, kn, k2, partT);
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      with_parts(ns ? partN : partT, [&](auto X) {
        X.rows = rows_;
        X.nsplit = ns ? sN3 : sT3;
        X.rb = rb_for(mode, k2);
        X.r_out = k2;
        X.r_in = rr;
        X.base = ns ? K2 : L2;
        X.ld_base = rows_;
        X.X = ns ? U : V;
        X.ldx = rows_;
        X.Y = ns ? A->G1 : A->G2;
        X.ldy = rr;
        X.out = ns ? P : Q;
        X.ldo = rows_;
        rows(X);
      });
    }
    // the completions' reverse, and Ub0 / Vb0 up to the last pass
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      double* acc = ns ? P : Q;                 // [:, :r] Ub0 so far, [:, r:2r] Qbar
      const double* Qbar = acc + rows_ * rr;
      const double* Qc = (ns ? Uhat : Vhat) + rows_ * rr;
      const double* B0 = ns ? U : V;
      const double* X1 = ns ? K : L;
      double* Zb = ns ? K2 : L2;
      double* tmp = Zb + rows_ * rcap;
      double* Xb1 = ns ? U2 : V2;               // Kbar1 / Lbar1
      {
        RowArgs<double> X;  // copyltu(-Qbar' Q)
        X.rows = rows_;
        X.r_out = rr;
        X.base = Qbar;
        X.ld_base = rows_;
        X.reduce = 1;
        X.W = Qc;
        X.ldw = rows_;
        X.rw = rr;
        X.epi.kind = EPI_COPYLTU;
        X.epi.dst = ns ? A->CqU : A->CqV;
        rows(X);
      }
      {
        RowArgs<double> X;  // tmp = Qbar + Q Cq
        X.rows = rows_;
        X.r_out = rr;
        X.r_in = rr;
        X.base = Qbar;
        X.ld_base = rows_;
        X.X = Qc;
        X.ldx = rows_;
        X.Y = ns ? A->CqU : A->CqV;
        X.ldy = rr;
        X.out = tmp;
        X.ldo = rows_;
        rows(X);
      }
      {
        RowArgs<double> X;  // Zbar = tmp R^{-T}; Zbar' B0
        X.rows = rows_;
        X.r_out = rr;
        X.r_in = rr;
        X.X = tmp;
        X.ldx = rows_;
        X.Y = ns ? A->RUinvT : A->RVinvT;
        X.ldy = rr;
        X.out = Zb;
        X.ldo = rows_;
        X.reduce = 1;
        X.W = B0;
        X.ldw = rows_;
        X.rw = rr;
        X.epi.kind = EPI_COPY;
        X.epi.dst = ns ? A->ZU : A->ZV;
        rows(X);
      }
      {
        RowArgs<double> X;  // K1' U0
        X.rows = rows_;
        X.r_out = rr;
        X.base = X1;
        X.ld_base = rows_;
        X.reduce = 1;
        X.W = B0;
        X.ldw = rows_;
        X.rw = rr;
        X.epi.kind = EPI_COPY;
        X.epi.dst = ns ? A->KU : A->LV;
        rows(X);
      }
      {
        RowArgs<double> X;  // Kbar1 = Zbar - U0 (Zbar'U0)', packed for the last pass
        X.rows = rows_;
        X.r_out = rr;
        X.r_in = rr;
        X.base = Zb;
        X.ld_base = rows_;
        X.X = B0;
        X.ldx = rows_;
        X.Y = ns ? A->ZU : A->ZV;
        X.ldy = rr;
        X.transY = 1;
        X.xy_sign = -1.0;
        X.out = Xb1;
        X.ldo = rows_;
        set_pack_outputs(X, !ns, mode, rb_for(mode, rr));  // Kbar1 for pass T, Lbar1 for pass N
        rows(X);
      }
      // acc[:, :r] += Xhat G3 - Zbar (K1'U0) - K1 (Zbar'U0) + Kbar1 S0' (V side: ... + Lbar1 S0)
      const double* Xs[4] = {ns ? Uhat : Vhat, Zb, X1, Xb1};
      const double* Ys[4] = {ns ? A->G3 : A->G4, ns ? A->KU : A->LV, ns ? A->ZU : A->ZV, S};
      const int ri[4] = {k2, rr, rr, rr};
      const int ldys[4] = {k2, rr, rr, (int)rcap};
      const int trs[4] = {0, 0, 0, ns ? 1 : 0};
      const double sg[4] = {1.0, -1.0, -1.0, 1.0};
      for (int u = 0; u < 4; ++u) {
        RowArgs<double> X;
        X.rows = rows_;
        X.r_out = rr;
        X.r_in = ri[u];
        X.base = acc;
        X.ld_base = rows_;
        X.X = Xs[u];
        X.ldx = rows_;
        X.Y = Ys[u];
        X.ldy = ldys[u];
        X.transY = trs[u];
        X.xy_sign = sg[u];
        X.out = acc;
        X.ldo = rows_;
        rows(X);
      }
    }
    const int sN4 = passN(kp, kn, rr, partN);  // dA Lbar1
    const int sT4 = passT(kp, kn, rr, partT);  // dA' Kbar1
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      with_parts(ns ? partN : partT, [&](auto X) {
        X.rows = ns ? n : m;
        X.nsplit = ns ? sN4 : sT4;
        X.rb = rb_for(mode, rr);
        X.r_out = rr;
        X.base = ns ? P : Q;
        X.ld_base = X.rows;
        X.out = ns ? P : Q;
        X.ldo = X.rows;
        rows(X);
      });
    }
    pkN_mode = pkT_mode = -1;  // the panels hold adjoint operands now
  }

  // SVD of the 2r x 2r core small->Shat, truncation, and the rotation of Uhat,
  // Vhat into the new state (rank_adaptive_unconventional.jl:222-231).
  // ra_finish in two halves, so that a batch of contexts can queue the SVDs of
  // all its members before waiting for any rank: ra_svd launches the SVD and
  // copies the new rank to pinned host memory; ra_rotate waits for it and
  // forms U, S, V of the new rank.
  int64_t ra_finish(int k2, double tol, int64_t rmax) {
    ra_svd(k2, tol, rmax);
    return ra_rotate();
  }
  void ra_svd(int k2, double tol, int64_t rmax) {
    SvdArgs sv;
    sv.S = small->Shat;
    sv.lds = k2;
    sv.k = k2;
    sv.U = small->Ush;
    sv.V = small->Vsh;
    sv.sigma = small->sigma;
    sv.r_out = &small->r_new;
    sv.tol = tol;
    // clipped to the caller's rmax and to min(n, m) (singular values past it
    // are those of dependent directions), not to the capacity: ra_rotate
    // reports a rank above the capacity instead of truncating to it
    sv.rmax = std::min<int64_t>(rmax, std::min(n, m));
    sv.work = svd_work;
    timed(LRI_KERNEL_SMALL, [&] { return launch_jacobi_svd(sv, stream); });
    if (!defer_rank_copy) HIPCHECK(hipMemcpyAsync(r_host, &small->r_new, sizeof(int), hipMemcpyDeviceToHost, stream));
    ra_k2 = k2;
  }

  // A rank-adaptive DE step up to its SVD as a captured graph (captured on first
  // use for the current rank and step parameters, replayed after), then the
  // copy of the new rank; ra_rotate completes the step.  The step's launches
  // read and write only buffers fixed at creation, so a graph stays valid
  // while the rank is unchanged.
  void de_bug_ra_queue_graph(double dt, int64_t nsub, double tol, int64_t rmax) {
    hipGraphExec_t exec = nullptr;
    for (const RaGraph& g : ra_graphs)
      if (g.r == r && g.nsub == nsub && g.rmax == rmax && g.dt == dt && g.tol == tol) exec = g.exec;
    if (!exec) {
      hipGraph_t graph = nullptr;
      HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      defer_rank_copy = true;
      try {
        de_bug_ra(dt, nsub, tol, rmax, false);
      } catch (...) {
        defer_rank_copy = false;
        (void)hipStreamEndCapture(stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      defer_rank_copy = false;
      HIPCHECK(hipStreamEndCapture(stream, &graph));
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHECK(e);
      if (ra_graphs.size() >= 8) {
        (void)hipGraphExecDestroy(ra_graphs.front().exec);
        ra_graphs.erase(ra_graphs.begin());
      }
      ra_graphs.push_back(RaGraph{r, nsub, rmax, dt, tol, exec});
    }
    HIPCHECK(hipGraphLaunch(exec, stream));
    ra_k2 = 2 * (int)r;
    HIPCHECK(hipMemcpyAsync(r_host, &small->r_new, sizeof(int), hipMemcpyDeviceToHost, stream));
  }
  // Captured graphs bake in the operator's buffers and its control flow
  // (analytic eigenbasis or not, the quadratic term): a new operator drops them.
  void drop_graphs() {
    ++epoch;
    HIPCHECK(hipStreamSynchronize(stream));
    for (auto& g : ra_graphs) (void)hipGraphExecDestroy(g.exec);
    ra_graphs.clear();
  }
  int64_t ra_rotate() {
    HIPCHECK(hipStreamSynchronize(stream));
    const int r1 = *r_host;
    ra_check_rank(r1);
    ra_apply(ra_k2, r1);
    r = r1;
    return r1;
  }
  void ra_check_rank(int r1) {
    if (r1 > rcap) {
      // the state is left at the step's start; the panels hold the augmented bases
      pkN_mode = pkT_mode = -1;
      throw Failure{LRI_ENOTSUP, "the rank-adaptive step needs rank " + std::to_string(r1) +
                                     " (truncation at tol within rmax), above the context capacity " +
                                     std::to_string(rcap) + " (create the context with a larger rcap, at most " +
                                     std::to_string(LRI_MAX_RANK) + ")"};
    }
    REQUIRE(r1 >= 1, "internal: rank-adaptive step returned an invalid rank");
  }
  // U = Uhat Ush[:, :r1], V = Vhat Vsh[:, :r1], S = diag(sigma[:r1]) (the
  // launches only: the caller sets r)
  void ra_apply(int k2, int r1) {
    for (int side = 0; side < 2; ++side) {
      RowArgs<double> A;
      A.rows = side == 0 ? n : m;
      A.r_out = r1;
      A.r_in = k2;
      A.X = side == 0 ? Uhat : Vhat;
      A.ldx = A.rows;
      A.Y = side == 0 ? small->Ush : small->Vsh;
      A.ldy = k2;
      A.out = side == 0 ? U : V;
      A.ldo = A.rows;
      set_pack_outputs(A, side == 1, pass_mode(), rb_for(pass_mode(), r1));
      rows(A);
    }
    const SetDiagArgs d{S, small->sigma, (int)rcap, (int)rcap, r1};
    if (!record_launch(KID_SET_DIAG, 0, 1, 256, 0, d)) {
      hipLaunchKernelGGL(k_set_diag, dim3(1), dim3(256), 0, stream, d);
      HIPCHECK(hipGetLastError());
    }
  }
  // dst[0..count) = src[0..count) on the stream (recorded while a recorder is installed)
  void dcopy(double* dst, const double* src, int64_t count) {
    if (count <= 0) return;
    const CopyArgs a{dst, src, count};
    const unsigned grid = (unsigned)std::min<int64_t>((count + 4 * kCopyThreads - 1) / (4 * kCopyThreads), 256);
    if (!record_launch(KID_COPY, 0, grid, kCopyThreads, 0, a))
      HIPCHECK(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToDevice, stream));
  }

  // ------------------------------------------------------------ matrix ODE
  // Galerkin data of an orthonormal basis X (rows x k) on side A (n rows,
  // nside) or B (m rows): Cg = X' T X (k x k) and FXT = F' X (q x k), with T the
  // side's tridiagonal operator and F its forcing factor (G or H).
  void de_galerkin(const double* X, bool nside, int k, double* Cg, double* FXT) {
    const int64_t rows_ = nside ? n : m;
    RowArgs<double> A;
    A.rows = rows_;
    A.r_out = k;
    A.base = X;
    A.ld_base = rows_;
    A.st_d = nside ? op_ad : op_bd;
    A.st_e = nside ? op_ae : op_be;
    A.st_periodic = nside ? op_aper : op_bper;
    A.reduce = 1;
    A.W = X;
    A.ldw = rows_;
    A.rw = k;
    A.epi.kind = EPI_COPY;
    A.epi.dst = Cg;
    rows(A);
    if (op_q > 0) {
      RowArgs<double> B;
      B.rows = rows_;
      B.r_out = op_q;
      B.base = nside ? op_G : op_H;
      B.ld_base = rows_;
      B.reduce = 1;
      B.W = X;
      B.ldw = rows_;
      B.rw = k;
      B.epi.kind = EPI_COPY;
      B.epi.dst = FXT;
      rows(B);
    }
  }
  SymEigArgs eig_args(int slot, const double* Cg, int k, double* W, double* ev) {
    SymEigArgs a;
    a.M = Cg;
    a.ldm = k;
    a.k = k;
    a.W = W;
    a.ev = ev;
    a.work = de->ework[slot];
    a.Uscr = de->Uscr[slot];
    a.svd_work = slot == 0 ? svd_work : de->svd2;
    a.r_scr = &de->rscr[slot];
    return a;
  }
  void de_eig(const double* Cg, int k, double* W, double* ev) {
    const SymEigArgs a = eig_args(0, Cg, k, W, ev);
    timed(LRI_KERNEL_SMALL, [&] { return launch_sym_eig(&a, 1, stream); });
  }
  // the two Galerkin eigenproblems of a step, one workgroup each in one launch
  void de_eig2(int ka, int kb) {
    const SymEigArgs a[2] = {eig_args(0, de->Ga, ka, de->Wa, de->alpha), eig_args(1, de->Gb, kb, de->Wb, de->beta)};
    timed(LRI_KERNEL_SMALL, [&] { return launch_sym_eig(a, 2, stream); });
  }
  // One K-type sub-problem on side `nside` (rows n, operator A, forcing G; or
  // rows m, operator B, forcing H):  Z' = T Z + Z Co + F FXo,  Z0 = X0 op(C0),
  // where Co = Wo diag(ev) Wo' is the other side's Galerkin matrix and FXo its
  // forcing projection (q x k).  de_prepare writes Zbar0 = Z0 Wo to `scratch`
  // (slot: which of the two concurrent sub-problems) and returns the RK4
  // arguments; de_run integrates one or two prepared sub-problems in one
  // launch; de_finish writes Z = Zbar Wo' to `out` with the reduction `epi` on
  // its Gram matrix (EPI_NONE: no reduction).
  RK4ColArgs de_prepare(const double* X0, bool nside, int k, const double* C0, int ldc, int transC,
                        const double* Wo, const double* ev, const double* FXo, int slot, double h, int64_t nsub,
                        double* scratch) {
    const int64_t rows_ = nside ? n : m;
    double* CW = de->CW[slot];
    double* coef = de->coef[slot];
    timed(LRI_KERNEL_SMALL, [&] { return launch_de_prep(C0, ldc, transC, Wo, FXo, k, op_q, CW, coef, stream); });
    RowArgs<double> A;  // Zbar0 = X0 op(C0) Wo
    A.rows = rows_;
    A.r_out = k;
    A.r_in = k;
    A.X = X0;
    A.ldx = rows_;
    A.Y = CW;
    A.ldy = k;
    A.out = scratch;
    A.ldo = rows_;
    rows(A);
    RK4ColArgs P;
    P.side = nside ? 0 : 1;
    P.n = rows_;
    P.ncols = k;
    P.X = scratch;
    P.ldx = rows_;
    P.Y = scratch;
    P.ldy = rows_;
    P.st_d = nside ? op_ad : op_bd;
    P.st_e = nside ? op_ae : op_be;
    P.periodic = nside ? op_aper : op_bper;
    P.mu = ev;
    P.g = nside ? op_G : op_H;
    P.ldg = rows_;
    P.coef = coef;
    P.q = op_q;
    P.h = h;
    P.nsub = nsub;
    return P;
  }
  // out (rn x k) = op(Qm) Bsrc for an rn x rn FP64 matrix Qm (op: transpose when
  // trans), on the FP64 streaming passes with a single operand: pack Bsrc into
  // the panel with rn rows, one pass, then the sum of the split partials.
  void spectral_gemm(bool trans, const double* Qm, int64_t rn, const double* Bsrc, int k, double* out) {
    const int rb = rb_for(PASS_F64, k);
    double* pk = (rn == n) ? pkT_f64 : pkN_f64;
    RowArgs<double> A;
    A.rows = rn;
    A.r_out = k;
    A.base = Bsrc;
    A.ld_base = rn;
    A.pk_f64 = pk;
    A.pk_cols = rb;
    rows(A);
    if (rn == n) pkT_mode = -1;  // the panel no longer holds a state basis
    else pkN_mode = -1;
    PassArgs a;
    a.anext = Qm;
    a.aprev = nullptr;
    a.dtype = LRI_F64;
    a.lda = rn;
    a.n = rn;
    a.m = rn;
    a.rb = rb;
    a.B = pk;
    a.part = partT;
    const int64_t need = plan_pass(PASS_F64, trans, rn, rn, rb, &a.splits, &a.chunk);
    REQUIRE((size_t)need * sizeof(double) <= part_bytes, "internal: partial buffer too small");
    timed(LRI_KERNEL_DE, [&] { return trans ? launch_passT(PASS_F64, a, stream) : launch_passN(PASS_F64, a, stream); });
    RowArgs<double> S;
    S.rows = rn;
    S.nsplit = a.splits;
    S.rb = rb;
    S.r_out = k;
    S.parts = (const double*)partT;
    S.out = out;
    S.ldo = rn;
    rows(S);
  }

  // Integrate one or two prepared sub-problems: sides with a constant operator
  // in the analytic eigenbasis (Zhat = Q' Z, closed-form flow, Z = Q Zhat, with
  // Zhat in the P / Q work buffers), the others in one k_rk4_columns launch.
  void de_run(const RK4ColArgs& P0, const RK4ColArgs& P1 = RK4ColArgs()) {
    if (resident) {
      // the state lives in the eigenbasis: the closed-form flow, no transforms
      for (const RK4ColArgs* Pp : {&P0, &P1}) {
        const RK4ColArgs& Pa = *Pp;
        if (Pa.ncols < 1) continue;
        REQUIRE(Pa.Y == Pa.X, "internal: resident flow in place");
        timed(LRI_KERNEL_DE, [&] {
          return launch_spectral_flow(Pa.Y, Pa.n, Pa.ncols, specLam[Pa.side], Pa.mu, specF[Pa.side], Pa.coef, Pa.q,
                                      Pa.h, Pa.nsub, stream);
        });
      }
      return;
    }
    if (P0.ncols > 0 && P1.ncols > 0 && spec[P0.side] && spec[P1.side] && specQ[P0.side] == specQ[P1.side] &&
        P0.n == P1.n && P1.X == P0.X + P0.n * P0.ncols && P1.Y == P1.X && P0.Y == P0.X &&
        P0.ncols + P1.ncols <= 64) {
      // one eigenbasis, adjacent columns: one pair of transforms over both
      const int k0 = P0.ncols, k1 = P1.ncols;
      double* Zh = this->P;
      spectral_gemm(true, specQ[P0.side], P0.n, P0.X, k0 + k1, Zh);
      timed(LRI_KERNEL_SMALL, [&] {
        hipError_t e = launch_spectral_flow(Zh, P0.n, k0, specLam[P0.side], P0.mu, specF[P0.side], P0.coef, P0.q,
                                            P0.h, P0.nsub, stream);
        if (e != hipSuccess) return e;
        return launch_spectral_flow(Zh + P0.n * k0, P1.n, k1, specLam[P1.side], P1.mu, specF[P1.side], P1.coef,
                                    P1.q, P1.h, P1.nsub, stream);
      });
      spectral_gemm(false, specQ[P0.side], P0.n, Zh, k0 + k1, P0.Y);
      return;
    }
    RK4ColArgs rk[2];
    int nrk = 0;
    for (const RK4ColArgs* Pp : {&P0, &P1}) {
      const RK4ColArgs& Pa = *Pp;
      if (Pa.ncols < 1) continue;
      if (!spec[Pa.side]) {
        rk[nrk++] = Pa;
        continue;
      }
      double* Zh = Pa.side == 0 ? this->P : this->Q;
      spectral_gemm(true, specQ[Pa.side], Pa.n, Pa.X, Pa.ncols, Zh);
      timed(LRI_KERNEL_SMALL, [&] {
        return launch_spectral_flow(Zh, Pa.n, Pa.ncols, specLam[Pa.side], Pa.mu, specF[Pa.side], Pa.coef, Pa.q,
                                    Pa.h, Pa.nsub, stream);
      });
      spectral_gemm(false, specQ[Pa.side], Pa.n, Zh, Pa.ncols, Pa.Y);
    }
    if (nrk > 0) {
      const RK4ColArgs none;
      timed(LRI_KERNEL_DE, [&] { return launch_rk4_columns(rk[0], nrk > 1 ? rk[1] : none, stream); });
    }
  }

  void de_finish(bool nside, int k, const double* Wo, double* out, const double* scratch, EpiArgs epi) {
    const int64_t rows_ = nside ? n : m;
    RowArgs<double> B;  // Z = Zbar Wo'
    B.rows = rows_;
    B.r_out = k;
    B.r_in = k;
    B.X = scratch;
    B.ldx = rows_;
    B.Y = Wo;
    B.ldy = k;
    B.transY = 1;
    B.out = out;
    B.ldo = rows_;
    if (epi.kind != EPI_NONE) {
      B.reduce = 1;
      B.epi = epi;
    }
    rows(B);
  }
  void de_substep(const double* X0, bool nside, int k, const double* C0, int ldc, int transC, const double* Wo,
                  const double* ev, const double* FXo, double h, int64_t nsub, double* out, double* scratch,
                  EpiArgs epi) {
    de_run(de_prepare(X0, nside, k, C0, ldc, transC, Wo, ev, FXo, 0, h, nsub, scratch));
    de_finish(nside, k, Wo, out, scratch, epi);
  }
  void de_sstep(int k, int l, const double* S0, int lds0, int transS0, double* S1, int lds1, double sgn, double h,
                int64_t nsub) {
    SStepArgs P;
    P.k = k;
    P.l = l;
    P.q = op_q;
    P.S0 = S0;
    P.lds0 = lds0;
    P.transS0 = transS0;
    P.S1 = S1;
    P.lds1 = lds1;
    P.Wa = de->Wa;
    P.alpha = de->alpha;
    P.Wb = de->Wb;
    P.beta = de->beta;
    P.gaT = de->gaT;
    P.hbT = de->hbT;
    P.sgn = sgn;
    P.h = h;
    P.nsub = nsub;
    P.work = de->swork;
    timed(LRI_KERNEL_SMALL, [&] { return launch_s_substep(P, stream); });
  }

  // primal_LT_step! with the DE sub-problems (projector_splitting.jl:52-80,129-152)
  // The QRs of the matrix-ODE sub-steps use shifted CholeskyQR3: a stiff or
  // strongly graded state (e.g. singular values down to 1e-8) makes the K and
  // L factors too ill-conditioned for CholeskyQR2.  A numerically dependent
  // direction (a stiff sub-step can collapse K onto a few directions) becomes
  // a zero column instead of an error; Householder QR would put an arbitrary
  // orthonormal direction there.
  static EpiArgs shifted_chol(int64_t rows_) {
    EpiArgs e;
    e.kind = EPI_CHOL1;
    e.shift_rows = rows_;
    e.breakdown_is_error = 0;
    return e;
  }
  static EpiArgs de_epi(int kind = EPI_NONE) {
    EpiArgs e;
    e.kind = kind;
    e.breakdown_is_error = 0;
    return e;
  }

  // ---------------------------------------------------- quadratic operator
  // T = V x V x V (m-side basis X, no stencil) or W = (D U) x U x U (n side)
  void quad_tensor(const double* X, bool nside, int k, double* out) {
    TensorArgs Q;
    Q.rows = nside ? n : m;
    Q.k = k;
    Q.Y = X;
    Q.ldy = Q.rows;
    if (nside) {
      Q.dl = op_dl;
      Q.dd = op_dd;
      Q.du = op_du;
      Q.dper = op_dper;
    }
    Q.part = quad_part;
    timed(LRI_KERNEL_SMALL, [&] { return launch_tensor3(Q, out, stream); });
  }
  // Galerkin matrices, forcing projections and the tensor of one basis
  void quad_galerkin(const double* X, bool nside, int k) {
    if (nside) {
      de_galerkin(X, true, k, de->Ga, de->gaT);
      quad_tensor(X, true, k, quad_W);
    } else {
      de_galerkin(X, false, k, de->Gb, de->hbT);
      quad_tensor(X, false, k, quad_T);
    }
  }
  // X0 = B0 op(C) (rows x k) into X, the initial value of a K or L sub-problem
  void quad_init(const double* B0, int64_t rows_, int k, const double* C, int ldc, int transC, double* X) {
    RowArgs<double> A;
    A.rows = rows_;
    A.r_out = k;
    A.r_in = k;
    A.X = B0;
    A.ldx = rows_;
    A.Y = C;
    A.ldy = ldc;
    A.transY = transC;
    A.out = X;
    A.ldo = rows_;
    rows(A);
  }
  // The stage arguments of the K (nside: operator A, D on K, forcing G) or L
  // sub-problem (operator B, no stencil, forcing H) with the Galerkin matrix M,
  // forcing projection hv and tensor T of the other side.
  QuadStageArgs quad_stage_args(bool nside, int k, const double* M, const double* hv, const double* T, double h) {
    const int64_t rows_ = nside ? n : m;
    QuadStageArgs Q;
    Q.n = rows_;
    Q.k = k;
    Q.ld = rows_;
    Q.h = h;
    Q.st_d = nside ? op_ad : op_bd;
    Q.st_e = nside ? op_ae : op_be;
    Q.periodic = nside ? op_aper : op_bper;
    Q.M = M;
    Q.g = nside ? op_G : op_H;
    Q.ldg = rows_;
    Q.q = op_q;
    Q.hv = hv;
    Q.T = T;
    if (nside) {
      Q.dl = op_dl;
      Q.dd = op_dd;
      Q.du = op_du;
      Q.dper = op_dper;
    }
    return Q;
  }
  // N RK4 steps of the K (nside: M = V'BV, forcing G (H'V), tensor T, D on K) or
  // L sub-problem (M = U'AU, forcing H (G'U), tensor W, no stencil) on X in place;
  // scr holds 2 k columns, acc k columns (ld rows).  With `store`, the start of
  // sub-step st is kept at store + st rows k.
  void quad_rows_ex(bool nside, int k, const double* M, const double* hv, const double* T, double* X, double* scr,
                    double* acc, double h, int64_t nsub, double* store) {
    const int64_t rows_ = nside ? n : m;
    QuadStageArgs Q = quad_stage_args(nside, k, M, hv, T, h);
    Q.X0 = X;
    Q.Acc = acc;
    double* xa = scr;
    double* xb = scr + rows_ * k;
    const double* in[4] = {X, xa, xb, xa};
    double* outp[4] = {xa, xb, xa, nullptr};
    for (int64_t st = 0; st < nsub; ++st) {
      if (store)
        HIPCHECK(hipMemcpyAsync(store + st * rows_ * k, X, sizeof(double) * rows_ * k, hipMemcpyDeviceToDevice,
                                stream));
      for (int s = 0; s < 4; ++s) {
        Q.stage = s;
        Q.Xs = in[s];
        Q.Xn = outp[s];
        timed(LRI_KERNEL_DE, [&] { return launch_quad_stage(Q, stream); });
      }
    }
  }
  void quad_rows(bool nside, int k, double* X, double* scr, double* acc, double h, int64_t nsub) {
    quad_rows_ex(nside, k, nside ? de->Gb : de->Ga, nside ? de->hbT : de->gaT, nside ? quad_T : quad_W, X, scr, acc,
                 h, nsub, nullptr);
  }
  // Gram of X (rows x k) and the shifted first CholeskyQR factor
  void quad_gram(const double* X, int64_t rows_, int k, EpiArgs epi) {
    RowArgs<double> A;
    A.rows = rows_;
    A.r_out = k;
    A.base = X;
    A.ld_base = rows_;
    A.reduce = 1;
    A.epi = epi;
    rows(A);
  }
  void quad_sstep(int k, int l, const double* S0, int lds0, int transS0, double* S1, int lds1, double sgn, double h,
                  int64_t nsub) {
    QuadSArgs Q;
    Q.k = k;
    Q.l = l;
    Q.q = op_q;
    Q.S0 = S0;
    Q.lds0 = lds0;
    Q.transS0 = transS0;
    Q.S1 = S1;
    Q.lds1 = lds1;
    Q.Ga = de->Ga;
    Q.Gb = de->Gb;
    Q.gaT = de->gaT;
    Q.hbT = de->hbT;
    Q.W = quad_W;
    Q.T = quad_T;
    Q.sgn = sgn;
    Q.h = h;
    Q.nsub = nsub;
    Q.work = quad_work;
    timed(LRI_KERNEL_SMALL, [&] { return launch_quad_sstep(Q, stream); });
  }

  void quad_ksl_primal(double dt, int64_t nsub) {
    const int rr = (int)r;
    const double h = dt / (double)nsub;
    const EpiArgs none = de_epi();
    quad_galerkin(V, false, rr);  // V0'BV0, H'V0, T(V0)
    quad_init(U, n, rr, S, (int)rcap, 0, K);
    quad_rows(true, rr, K, K2, P, h, nsub);
    quad_gram(K, n, rr, shifted_chol(n));
    cholqr_tail(K, K2, n, rr, U, false, nullptr, 0, none, 0, 3);  // U1, R
    quad_galerkin(U, true, rr);                                    // U1'AU1, G'U1, W(U1)
    quad_sstep(rr, rr, small->Rtot, rr, 0, de->Snew, rr, -1.0, h, nsub);  // S~
    quad_init(V, m, rr, de->Snew, rr, 1, L);
    quad_rows(false, rr, L, L2, Q, h, nsub);
    quad_gram(L, m, rr, shifted_chol(m));
    cholqr_tail(L, L2, m, rr, V, true, nullptr, 0, none, 2, 3);  // V1, S = R'
  }

  void quad_ksl_dual(double dt, int64_t nsub) {
    const int rr = (int)r;
    const double h = dt / (double)nsub;
    const EpiArgs none = de_epi();
    quad_galerkin(U, true, rr);  // U0'AU0, G'U0, W(U0)
    quad_init(V, m, rr, S, (int)rcap, 1, L);
    quad_rows(false, rr, L, L2, Q, h, nsub);
    quad_gram(L, m, rr, shifted_chol(m));
    cholqr_tail(L, L2, m, rr, V, true, nullptr, 0, none, 0, 3);  // V1, R
    quad_galerkin(V, false, rr);                                  // V1'BV1, H'V1, T(V1)
    quad_sstep(rr, rr, small->Rtot, rr, 1, de->Snew, rr, -1.0, h, nsub);  // S~ from R'
    quad_init(U, n, rr, de->Snew, rr, 0, K);
    quad_rows(true, rr, K, K2, P, h, nsub);
    quad_gram(K, n, rr, shifted_chol(n));
    cholqr_tail(K, K2, n, rr, U, false, nullptr, 0, none, 1, 3);  // U1, S = R
  }

  void quad_bug(double dt, int64_t nsub) {
    const int rr = (int)r;
    const double h = dt / (double)nsub;
    if (!de_gal_fresh) {
      quad_galerkin(V, false, rr);
      quad_galerkin(U, true, rr);
    }
    de_gal_fresh = false;
    quad_init(U, n, rr, S, (int)rcap, 0, K);
    quad_rows(true, rr, K, K2, P, h, nsub);
    quad_init(V, m, rr, S, (int)rcap, 1, L);
    quad_rows(false, rr, L, L2, Q, h, nsub);
    quad_gram(K, n, rr, shifted_chol(n));
    cholqr_tail(K, K2, n, rr, U2, false, U, rr, de_epi(EPI_BUG_M), 0, 3);  // U1, M = U1'U0
    quad_gram(L, m, rr, shifted_chol(m));
    EpiArgs eN = de_epi(EPI_BUG_N);  // N = V1'V0, Sinit = M S0 N'
    eN.S = S;
    eN.ld_s = (int)rcap;
    cholqr_tail(L, L2, m, rr, V2, true, V, rr, eN, 0, 3);
    quad_galerkin(U2, true, rr);
    quad_galerkin(V2, false, rr);
    quad_sstep(rr, rr, small->Sinit, rr, 0, S, (int)rcap, 1.0, h, nsub);
    std::swap(U, U2);
    std::swap(V, V2);
    de_gal_fresh = true;  // the Galerkin data and tensors now belong to U, V
  }

  // ------------------------------------------------- matrix-ODE adjoint
  // Reverse of quad_bug (unconventional_step! with the DE sub-problems,
  // unconventional.jl:133-157, :51-79) for the cotangents (Ub, Sb, Vb) of the
  // result (U1, S1, V1) in P[:, :r], adj->Sb1 and Q[:, :r]; leaves those of
  // (U0, S0, V0) in P[:, :r], adj->Sb0 and Q[:, :r], and in qadj->scal the
  // derivatives along scales a, b of the linear parts (F = a A Y + b Y B' + ...).
  // The forward step is recomputed keeping the RK4 sub-step starts of K and L;
  // then, with M = U1'U0, N = V1'V0 and the S sub-step from M S0 N':
  //   S reverse -> Sibar and the cotangents of U1'AU1, V1'BV1, G'U1, H'V1, W(U1), T(V1);
  //   Mb = Sibar N S0', Nb = Sibar' M S0, Sb0 = M' Sibar N;
  //   Ub1 += (basis cotangent of U1's Galerkin data) + U0 Mb', Vb1 likewise;
  //   Kbar1 = (Ub1 + U1 copyltu(-Ub1'U1)) RK^{-T}, Lbar1 likewise;
  //   K reverse -> Kbar0 and the cotangents of V0'BV0, H'V0, T(V0); L reverse likewise;
  //   Ub0 = (basis cotangent of U0's Galerkin data) + U1 Mb + Kbar0 S0',
  //   Vb0 = (... of V0's) + V1 Nb + Lbar0 S0,  Sb0 += U0'Kbar0 + Lbar0'V0.
  // ks: the S sub-step's size (r; 2r for the rank-adaptive step)
  size_t qadj_rows_need(int64_t nsub, int ks) const {
    const int rr = (int)r;
    const int64_t mx = std::max(n, m);
    const int plen = quad_adj_part_len(rr, op_q);
    return (size_t)(nsub * (n + m) * rr + 8 * mx * rr) +
           (size_t)plen * (quad_adj_grid(n, rr) + quad_adj_grid(m, rr)) + (size_t)quad_sadj_work(ks, ks, nsub);
  }
  void qadj_reserve(int64_t nsub, int ks = 0) {
    if (ks < 1) ks = (int)r;
    if (!qadj) {
      HIPCHECK(hipMalloc((void**)&qadj, sizeof(QuadAdj)));
      HIPCHECK(hipMemset(qadj, 0, sizeof(QuadAdj)));
    }
    const size_t need = qadj_rows_need(nsub, ks);
    if (need > qadj_rows_len) {
      HIPCHECK(hipStreamSynchronize(stream));
      if (qadj_rows) HIPCHECK(hipFree(qadj_rows));
      qadj_rows = nullptr;
      qadj_rows_len = 0;
      HIPCHECK(hipMalloc((void**)&qadj_rows, sizeof(double) * need));
      qadj_rows_len = need;
    }
  }
  // Reverse of the N RK4 steps of quad_rows_ex: store holds the sub-step starts,
  // ybar the cotangent of the result on entry and of the start on exit; the
  // summed parameter cotangents go to sum.  work: 7 rows x k.
  void quad_rows_reverse(bool nside, int k, const double* M, const double* hv, const double* T, const double* store,
                         double* ybar, double* work, double* part, double* sum, double h, int64_t nsub) {
    const int64_t rows_ = nside ? n : m;
    const int64_t sz = rows_ * k;
    double* x2 = work;
    double* x3 = x2 + sz;
    double* x4 = x3 + sz;
    double* xa = x4 + sz;
    double* xb = xa + sz;
    double* ya = xb + sz;
    double* accs = ya + sz;
    const int plen = quad_adj_part_len(k, op_q);
    const int grid = quad_adj_grid(rows_, k);
    HIPCHECK(hipMemsetAsync(part, 0, sizeof(double) * plen * grid, stream));
    QuadStageArgs F = quad_stage_args(nside, k, M, hv, T, h);
    F.Acc = accs;
    QuadAdjStageArgs R;
    R.n = rows_;
    R.k = k;
    R.ld = rows_;
    R.h = h;
    R.st_d = F.st_d;
    R.st_e = F.st_e;
    R.periodic = F.periodic;
    R.M = M;
    R.g = F.g;
    R.ldg = F.ldg;
    R.q = F.q;
    R.T = T;
    R.dl = F.dl;
    R.dd = F.dd;
    R.du = F.du;
    R.dper = F.dper;
    R.part = part;
    double* yb = ybar;
    double* yn = ya;
    for (int64_t st = nsub - 1; st >= 0; --st) {
      double* y = const_cast<double*>(store) + st * sz;  // read-only for stages 0..2
      F.X0 = y;
      const double* fin[3] = {y, x2, x3};
      double* fout[3] = {x2, x3, x4};
      for (int s = 0; s < 3; ++s) {  // the stage inputs X2, X3, X4 of this sub-step
        F.stage = s;
        F.Xs = fin[s];
        F.Xn = fout[s];
        timed(LRI_KERNEL_DE, [&] { return launch_quad_stage(F, stream); });
      }
      const double* in[4] = {y, x2, x3, x4};
      double* xbo = xa;
      const double* xbp = nullptr;
      for (int s = 3; s >= 0; --s) {
        R.stage = s;
        R.X = in[s];
        R.Ybar = yb;
        R.Xbp = xbp;
        R.Xbo = xbo;
        R.Acc = yn;
        timed(LRI_KERNEL_DE, [&] { return launch_quad_adj_stage(R, stream); });
        xbp = xbo;
        xbo = xbo == xa ? xb : xa;
      }
      std::swap(yb, yn);
    }
    if (yb != ybar) HIPCHECK(hipMemcpyAsync(ybar, yb, sizeof(double) * sz, hipMemcpyDeviceToDevice, stream));
    timed(LRI_KERNEL_SMALL, [&] { return launch_part_sum(part, grid, plen, sum, stream); });
  }
  BasisVjpArgs basis_vjp_args(bool nside, int k, const double* X, const double* Tb, const double* Gb,
                              const double* fb) {
    BasisVjpArgs B;
    B.rows = nside ? n : m;
    B.k = k;
    B.X = X;
    B.ldx = B.rows;
    if (nside) {
      B.dl = op_dl;
      B.dd = op_dd;
      B.du = op_du;
      B.dper = op_dper;
    }
    B.Tb = Tb;
    B.st_d = nside ? op_ad : op_bd;
    B.st_e = nside ? op_ae : op_be;
    B.periodic = nside ? op_aper : op_bper;
    B.Gb = Gb;
    B.F = nside ? op_G : op_H;
    B.ldf = B.rows;
    B.q = op_q;
    B.fb = fb;
    return B;
  }
  void quad_bug_adjoint(double dt, int64_t nsub) {
    const int rr = (int)r;
    const double h = dt / (double)nsub;
    const size_t d = sizeof(double);
    const int64_t mx = std::max(n, m);
    const int plen = quad_adj_part_len(rr, op_q);
    qadj_reserve(nsub);
    double* stK = qadj_rows;
    double* stL = stK + nsub * n * rr;
    double* work = stL + nsub * m * rr;  // 8 mx rr
    double* partK = work + 8 * mx * rr;
    double* partL = partK + (size_t)plen * quad_adj_grid(n, rr);
    double* swork = partL + (size_t)plen * quad_adj_grid(m, rr);
    double* kbar = work + 7 * mx * rr;  // the K (then L) cotangent, beside the 7-column reverse work
    QuadAdj* qa = qadj;
    // forward: Galerkin data of U0, V0 (kept), K and L with their sub-step starts
    quad_galerkin(V, false, rr);
    quad_galerkin(U, true, rr);
    de_gal_fresh = false;
    const size_t rsz = d * rr * rr, fsz = d * op_q * rr, tsz = d * rr * rr * rr;
    HIPCHECK(hipMemcpyAsync(qa->Ga0, de->Ga, rsz, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(qa->Gb0, de->Gb, rsz, hipMemcpyDeviceToDevice, stream));
    if (op_q) {
      HIPCHECK(hipMemcpyAsync(qa->gaT0, de->gaT, fsz, hipMemcpyDeviceToDevice, stream));
      HIPCHECK(hipMemcpyAsync(qa->hbT0, de->hbT, fsz, hipMemcpyDeviceToDevice, stream));
    }
    HIPCHECK(hipMemcpyAsync(qa->W0, quad_W, tsz, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(qa->T0, quad_T, tsz, hipMemcpyDeviceToDevice, stream));
    quad_init(U, n, rr, S, (int)rcap, 0, K);
    quad_rows_ex(true, rr, qa->Gb0, qa->hbT0, qa->T0, K, work, work + 2 * mx * rr, h, nsub, stK);
    quad_init(V, m, rr, S, (int)rcap, 1, L);
    quad_rows_ex(false, rr, qa->Ga0, qa->gaT0, qa->W0, L, work, work + 2 * mx * rr, h, nsub, stL);
    quad_gram(K, n, rr, shifted_chol(n));
    cholqr_tail(K, K2, n, rr, U2, false, U, rr, de_epi(EPI_BUG_M), 0, 3);  // U1, M = U1'U0
    HIPCHECK(hipMemcpyAsync(adj->RK, small->Rtot, rsz, hipMemcpyDeviceToDevice, stream));
    quad_gram(L, m, rr, shifted_chol(m));
    EpiArgs eN = de_epi(EPI_BUG_N);  // N = V1'V0, Sinit = M S0 N'
    eN.S = S;
    eN.ld_s = (int)rcap;
    cholqr_tail(L, L2, m, rr, V2, true, V, rr, eN, 0, 3);
    HIPCHECK(hipMemcpyAsync(adj->RL, small->Rtot, rsz, hipMemcpyDeviceToDevice, stream));
    quad_galerkin(U2, true, rr);  // U1'AU1, G'U1, W(U1)
    quad_galerkin(V2, false, rr);  // V1'BV1, H'V1, T(V1)
    // S sub-step reverse
    QuadSAdjArgs SA;
    SA.k = SA.l = rr;
    SA.q = op_q;
    SA.S0 = small->Sinit;
    SA.lds0 = rr;
    SA.Sb1 = adj->Sb1;
    SA.ldsb1 = rr;
    SA.Sb0 = qa->Sib;
    SA.Ga = de->Ga;
    SA.Gb = de->Gb;
    SA.gaT = de->gaT;
    SA.hbT = de->hbT;
    SA.W = quad_W;
    SA.T = quad_T;
    SA.sgn = 1.0;
    SA.h = h;
    SA.nsub = nsub;
    SA.Gab = qa->Gab1;
    SA.Gbb = qa->Gbb1;
    SA.gaTb = qa->gaTb1;
    SA.hbTb = qa->hbTb1;
    SA.Wb = qa->Wb1;
    SA.Tb = qa->Tb1;
    SA.work = swork;
    timed(LRI_KERNEL_SMALL, [&] { return launch_quad_sadj(SA, stream); });
    HIPCHECK(hipMemcpyAsync(adj->Sb1, qa->Sib, rsz, hipMemcpyDeviceToDevice, stream));
    // Mb, Nb, Sb0 = M' Sibar N, RK^{-T}, RL^{-T}
    timed(LRI_KERNEL_SMALL, [&] { return launch_bug_adj_pre(adj, small->M, small->N, S, (int)rcap, rr, stream); });
    double* Ubf = P + n * rr;
    double* Vbf = Q + m * rr;
    for (int side = 0; side < 2; ++side) {  // the full cotangents of U1 and V1
      const bool ns = side == 0;
      BasisVjpArgs B = basis_vjp_args(ns, rr, ns ? U2 : V2, ns ? qa->Wb1 : qa->Tb1, ns ? qa->Gab1 : qa->Gbb1,
                                      ns ? qa->gaTb1 : qa->hbTb1);
      B.X2 = ns ? U : V;
      B.ldx2 = B.rows;
      B.Y2 = ns ? adj->Mb : adj->Nb;
      B.ldy2 = rr;
      B.transY2 = 1;
      B.base = ns ? P : Q;
      B.ldb = B.rows;
      B.out = ns ? Ubf : Vbf;
      B.ldo = B.rows;
      timed(LRI_KERNEL_DE, [&] { return launch_basis_vjp(B, stream); });
    }
    // reverse of each sub-problem: QR backward, then the RK4 steps
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      RowArgs<double> A;  // copyltu(-Xbf' X1)
      A.rows = rows_;
      A.r_out = rr;
      A.base = ns ? Ubf : Vbf;
      A.ld_base = rows_;
      A.reduce = 1;
      A.W = ns ? U2 : V2;
      A.ldw = rows_;
      A.rw = rr;
      A.epi.kind = EPI_COPYLTU;
      A.epi.dst = ns ? adj->CqK : adj->CqL;
      rows(A);
      double* Z = ns ? K2 : L2;
      RowArgs<double> B;  // Z = Xbf + X1 Cq
      B.rows = rows_;
      B.r_out = rr;
      B.r_in = rr;
      B.base = ns ? Ubf : Vbf;
      B.ld_base = rows_;
      B.X = ns ? U2 : V2;
      B.ldx = rows_;
      B.Y = ns ? adj->CqK : adj->CqL;
      B.ldy = rr;
      B.out = Z;
      B.ldo = rows_;
      rows(B);
      RowArgs<double> C;  // Kbar1 = Z RK^{-T}
      C.rows = rows_;
      C.r_out = rr;
      C.r_in = rr;
      C.X = Z;
      C.ldx = rows_;
      C.Y = ns ? adj->RKinvT : adj->RLinvT;
      C.ldy = rr;
      C.out = kbar;
      C.ldo = rows_;
      rows(C);
      if (ns)
        quad_rows_reverse(true, rr, qa->Gb0, qa->hbT0, qa->T0, stK, kbar, work, partK, qa->sumK, h, nsub);
      else
        quad_rows_reverse(false, rr, qa->Ga0, qa->gaT0, qa->W0, stL, kbar, work, partL, qa->sumL, h, nsub);
      // Kbar0 -> K (Lbar0 -> L), and its reduction against U0 (V0): CK = Kbar0' U0, CL = Lbar0' V0
      RowArgs<double> D;
      D.rows = rows_;
      D.r_out = rr;
      D.base = kbar;
      D.ld_base = rows_;
      D.out = ns ? K : L;
      D.ldo = rows_;
      D.reduce = 1;
      D.W = ns ? U : V;
      D.ldw = rows_;
      D.rw = rr;
      D.epi.kind = EPI_COPY;
      D.epi.dst = ns ? adj->CK : adj->CL;
      rows(D);
    }
    // Ub0, Vb0
    const int k3 = rr * rr * rr;
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const double* sum = ns ? qa->sumL : qa->sumK;  // U0's data enter the L sub-problem, V0's the K one
      BasisVjpArgs B = basis_vjp_args(ns, rr, ns ? U : V, sum, sum + k3, sum + k3 + rr * rr);
      B.X2 = ns ? U2 : V2;
      B.ldx2 = B.rows;
      B.Y2 = ns ? adj->Mb : adj->Nb;
      B.ldy2 = rr;
      B.X3 = ns ? K : L;
      B.ldx3 = B.rows;
      B.Y3 = S;
      B.ldy3 = (int)rcap;
      B.transY3 = ns;  // Kbar0 S0', Lbar0 S0
      B.out = ns ? P : Q;
      B.ldo = B.rows;
      timed(LRI_KERNEL_DE, [&] { return launch_basis_vjp(B, stream); });
    }
    timed(LRI_KERNEL_SMALL, [&] { return launch_bug_adj_post(adj, rr, stream); });  // Sb0 += CK' + CL
    timed(LRI_KERNEL_SMALL, [&] {
      return launch_quad_adj_scales(qa->sumK, qa->sumL, plen, qa->sumL + k3, qa->Ga0, qa->Gab1, de->Ga,
                                    qa->sumK + k3, qa->Gb0, qa->Gbb1, de->Gb, rr, rr, qa->scal, stream);
    });
  }

  // Reverse of quad_ksl_primal (primal_LT_step! with the DE sub-problems,
  // projector_splitting.jl:52-80,129-152): K -> (U1, R1); S~ from R1 backwards in
  // time (sgn -1) with U1's and V0's Galerkin data; L from V0 S~' with U1's ->
  // (V1, R2); S1 = R2'.  Both QR factors carry cotangents:
  //   Lbar1 = (Vb1 + V1 copyltu(R2 Sb1 - Vb1'V1)) R2^{-T}  (R2bar = Sb1'),
  //   L reverse -> Lbar0, cotangents of U1's data;  Vb0 = Lbar0 S~ + ...,  S~bar = Lbar0' V0,
  //   S reverse -> R1bar and cotangents of U1's and V0's data,
  //   Kbar1 = (Ub1 + ... + U1 copyltu(R1 R1bar' - Ub1'U1)) R1^{-T},  K reverse -> Kbar0,
  //   Ub0 = Kbar0 S0',  Sb0 = U0'Kbar0,  Vb0 += (basis cotangent of V0's data).
  // Cotangents as in quad_bug_adjoint (P, adj->Sb1 / adj->Sb0, Q, qadj->scal).
  void quad_ksl_adjoint(double dt, int64_t nsub) {
    const int rr = (int)r;
    const double h = dt / (double)nsub;
    const size_t d = sizeof(double);
    const int64_t mx = std::max(n, m);
    const int plen = quad_adj_part_len(rr, op_q);
    qadj_reserve(nsub);
    double* stK = qadj_rows;
    double* stL = stK + nsub * n * rr;
    double* work = stL + nsub * m * rr;
    double* partK = work + 8 * mx * rr;
    double* partL = partK + (size_t)plen * quad_adj_grid(n, rr);
    double* swork = partL + (size_t)plen * quad_adj_grid(m, rr);
    double* kbar = work + 7 * mx * rr;
    QuadAdj* qa = qadj;
    const size_t rsz = d * rr * rr, fsz = d * op_q * rr, tsz = d * rr * rr * rr;
    // forward: V0's Galerkin data (kept in qa->Gb0, hbT0, T0), K with its sub-step starts
    quad_galerkin(V, false, rr);
    de_gal_fresh = false;
    HIPCHECK(hipMemcpyAsync(qa->Gb0, de->Gb, rsz, hipMemcpyDeviceToDevice, stream));
    if (op_q) HIPCHECK(hipMemcpyAsync(qa->hbT0, de->hbT, fsz, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(qa->T0, quad_T, tsz, hipMemcpyDeviceToDevice, stream));
    quad_init(U, n, rr, S, (int)rcap, 0, K);
    quad_rows_ex(true, rr, qa->Gb0, qa->hbT0, qa->T0, K, work, work + 2 * mx * rr, h, nsub, stK);
    quad_gram(K, n, rr, shifted_chol(n));
    cholqr_tail(K, K2, n, rr, U2, false, nullptr, 0, de_epi(), 0, 3);  // U1, R1
    HIPCHECK(hipMemcpyAsync(adj->RK, small->Rtot, rsz, hipMemcpyDeviceToDevice, stream));
    // U1's Galerkin data (kept in qa->Ga0, gaT0, W0); S~ from R1; L with its sub-step starts
    quad_galerkin(U2, true, rr);
    HIPCHECK(hipMemcpyAsync(qa->Ga0, de->Ga, rsz, hipMemcpyDeviceToDevice, stream));
    if (op_q) HIPCHECK(hipMemcpyAsync(qa->gaT0, de->gaT, fsz, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(qa->W0, quad_W, tsz, hipMemcpyDeviceToDevice, stream));
    quad_sstep(rr, rr, adj->RK, rr, 0, de->Snew, rr, -1.0, h, nsub);  // S~
    quad_init(V, m, rr, de->Snew, rr, 1, L);
    quad_rows_ex(false, rr, qa->Ga0, qa->gaT0, qa->W0, L, work, work + 2 * mx * rr, h, nsub, stL);
    quad_gram(L, m, rr, shifted_chol(m));
    cholqr_tail(L, L2, m, rr, V2, true, nullptr, 0, de_epi(), 0, 3);  // V1, R2
    HIPCHECK(hipMemcpyAsync(adj->RL, small->Rtot, rsz, hipMemcpyDeviceToDevice, stream));
    // Lbar1: R2 R2bar' = RL Sb1 (adj->Mb), then the QR rule
    timed(LRI_KERNEL_SMALL, [&] { return launch_ksl_adj_small(adj, rr, 0, stream); });
    auto qr_rule = [&](bool ns, const double* Xb, const double* Q1, const double* RRbar, double* Cq,
                       const double* RinvT, double* Z) {
      const int64_t rows_ = ns ? n : m;
      RowArgs<double> A;  // Cq = copyltu(R Rbar' - Xb' Q1)
      A.rows = rows_;
      A.r_out = rr;
      A.base = Xb;
      A.ld_base = rows_;
      A.reduce = 1;
      A.W = Q1;
      A.ldw = rows_;
      A.rw = rr;
      A.epi.kind = EPI_COPYLTU;
      A.epi.S = const_cast<double*>(RRbar);
      A.epi.ld_s = rr;
      A.epi.dst = Cq;
      rows(A);
      RowArgs<double> B;  // Z = Xb + Q1 Cq
      B.rows = rows_;
      B.r_out = rr;
      B.r_in = rr;
      B.base = Xb;
      B.ld_base = rows_;
      B.X = Q1;
      B.ldx = rows_;
      B.Y = Cq;
      B.ldy = rr;
      B.out = Z;
      B.ldo = rows_;
      rows(B);
      RowArgs<double> C;  // kbar = Z R^{-T}
      C.rows = rows_;
      C.r_out = rr;
      C.r_in = rr;
      C.X = Z;
      C.ldx = rows_;
      C.Y = RinvT;
      C.ldy = rr;
      C.out = kbar;
      C.ldo = rows_;
      rows(C);
    };
    qr_rule(false, Q, V2, adj->Mb, adj->CqL, adj->RLinvT, L2);
    quad_rows_reverse(false, rr, qa->Ga0, qa->gaT0, qa->W0, stL, kbar, work, partL, qa->sumL, h, nsub);  // Lbar0
    {
      RowArgs<double> A;  // Lbar0 -> L, S~bar = Lbar0' V0 -> qa->Sib
      A.rows = m;
      A.r_out = rr;
      A.base = kbar;
      A.ld_base = m;
      A.out = L;
      A.ldo = m;
      A.reduce = 1;
      A.W = V;
      A.ldw = m;
      A.rw = rr;
      A.epi.kind = EPI_COPY;
      A.epi.dst = qa->Sib;
      rows(A);
    }
    // S~ reverse (sgn -1 from R1): R1bar -> adj->CL
    QuadSAdjArgs SA;
    SA.k = SA.l = rr;
    SA.q = op_q;
    SA.S0 = adj->RK;
    SA.lds0 = rr;
    SA.Sb1 = qa->Sib;
    SA.ldsb1 = rr;
    SA.Sb0 = adj->CL;
    SA.Ga = qa->Ga0;
    SA.Gb = qa->Gb0;
    SA.gaT = qa->gaT0;
    SA.hbT = qa->hbT0;
    SA.W = qa->W0;
    SA.T = qa->T0;
    SA.sgn = -1.0;
    SA.h = h;
    SA.nsub = nsub;
    SA.Gab = qa->Gab1;
    SA.Gbb = qa->Gbb1;
    SA.gaTb = qa->gaTb1;
    SA.hbTb = qa->hbTb1;
    SA.Wb = qa->Wb1;
    SA.Tb = qa->Tb1;
    SA.work = swork;
    timed(LRI_KERNEL_SMALL, [&] { return launch_quad_sadj(SA, stream); });
    timed(LRI_KERNEL_SMALL, [&] { return launch_ksl_adj_small(adj, rr, 1, stream); });  // Nb = RK CL' = R1 R1bar'
    // the full cotangent of U1: Ub1 + its Galerkin data's (from the L and S reverses)
    const int k3 = rr * rr * rr;
    double* Ubf = P + n * rr;
    {
      BasisVjpArgs B = basis_vjp_args(true, rr, U2, qa->sumL, qa->sumL + k3, qa->sumL + k3 + rr * rr);
      B.Tb2 = qa->Wb1;
      B.Gb2 = qa->Gab1;
      B.fb2 = op_q ? qa->gaTb1 : nullptr;
      B.base = P;
      B.ldb = n;
      B.out = Ubf;
      B.ldo = n;
      timed(LRI_KERNEL_DE, [&] { return launch_basis_vjp(B, stream); });
    }
    qr_rule(true, Ubf, U2, adj->Nb, adj->CqK, adj->RKinvT, K2);
    quad_rows_reverse(true, rr, qa->Gb0, qa->hbT0, qa->T0, stK, kbar, work, partK, qa->sumK, h, nsub);  // Kbar0
    {
      RowArgs<double> A;  // Ub0 = Kbar0 S0'; Sb0 = U0' Kbar0 (C = Kbar0' U0, stored transposed)
      A.rows = n;
      A.r_out = rr;
      A.r_in = rr;
      A.X = kbar;
      A.ldx = n;
      A.Y = S;
      A.ldy = (int)rcap;
      A.transY = 1;
      A.out = P;
      A.ldo = n;
      rows(A);
      RowArgs<double> B;
      B.rows = n;
      B.r_out = rr;
      B.base = kbar;
      B.ld_base = n;
      B.reduce = 1;
      B.W = U;
      B.ldw = n;
      B.rw = rr;
      B.epi.kind = EPI_COPY;
      B.epi.transpose = 1;
      B.epi.dst = adj->Sb0;
      rows(B);
    }
    {
      BasisVjpArgs B = basis_vjp_args(false, rr, V, qa->sumK, qa->sumK + k3, qa->sumK + k3 + rr * rr);
      B.Tb2 = qa->Tb1;
      B.Gb2 = qa->Gbb1;
      B.fb2 = op_q ? qa->hbTb1 : nullptr;
      B.X3 = L;  // + Lbar0 S~
      B.ldx3 = m;
      B.Y3 = de->Snew;
      B.ldy3 = rr;
      B.out = Q;
      B.ldo = m;
      timed(LRI_KERNEL_DE, [&] { return launch_basis_vjp(B, stream); });
    }
    timed(LRI_KERNEL_SMALL, [&] {
      return launch_quad_adj_scales(qa->sumK, qa->sumL, plen, qa->sumL + k3, qa->Ga0, qa->Gab1, qa->Ga0,
                                    qa->sumK + k3, qa->Gb0, qa->Gbb1, qa->Gb0, rr, rr, qa->scal, stream);
    });
  }

  int64_t quad_bug_ra(double dt, int64_t nsub, double tol, int64_t rmax, bool complete = true) {
    const int rr = (int)r;
    const int k2 = 2 * rr;
    const double h = dt / (double)nsub;
    quad_galerkin(V, false, rr);
    quad_galerkin(U, true, rr);
    quad_init(U, n, rr, S, (int)rcap, 0, K);
    quad_rows(true, rr, K, K2, P, h, nsub);
    quad_init(V, m, rr, S, (int)rcap, 1, L);
    quad_rows(false, rr, L, L2, Q, h, nsub);
    HIPCHECK(hipMemcpyAsync(Uhat, U, sizeof(double) * n * rr, hipMemcpyDeviceToDevice, stream));
    EpiArgs eM;
    eM.kind = EPI_RA_M;
    complete_basis(K, U, n, rr, K2, K2 + n * rcap, Uhat + n * rr, eM);
    HIPCHECK(hipMemcpyAsync(Vhat, V, sizeof(double) * m * rr, hipMemcpyDeviceToDevice, stream));
    EpiArgs eN;
    eN.kind = EPI_RA_N;  // Sinit = M S0 N' (2r x 2r)
    eN.S = S;
    eN.ld_s = (int)rcap;
    complete_basis(L, V, m, rr, L2, L2 + m * rcap, Vhat + m * rr, eN);
    quad_galerkin(Uhat, true, k2);
    quad_galerkin(Vhat, false, k2);
    quad_sstep(k2, k2, small->Sinit, k2, 0, small->Shat, k2, 1.0, h, nsub);
    ra_svd(k2, tol, rmax);
    return complete ? ra_rotate() : 0;
  }

  // Reverse of quad_bug_ra (rankadaptive_unconventional_step! with the DE
  // sub-problems, rank_adaptive_unconventional.jl:59-86,194-233), the new rank r1
  // held fixed: the cotangents (Ub1, Sb1, Vb1) of the result (n x r1; r1 x r1, of
  // which the diagonal enters; m x r1) in P, radj->Sb1 and Q; leaves those of
  // (U0, S0, V0) in P[:, :r], adj->Sb0 and Q[:, :r], and in qadj->scal the
  // derivatives along the scales a, b of the linear parts.  The forward step is
  // recomputed up to the SVD of the 2r x 2r core, keeping the K and L sub-step
  // starts; with Uhat = [U0 | Q_U], Q_U R_U = (I - U0 U0') K1 (V likewise),
  // M = Uhat'U0, N = Vhat'V0 and the S sub-step from M S0 N':
  //   SVD rule -> Shbar (radj->core), Uhbar = Ub1 Us[:, :r1]';
  //   S reverse (2r) -> Sibar and the cotangents of Uhat's and Vhat's Galerkin data;
  //   Mb = Sibar N S0' (G3), Nb = Sibar' M S0 (G4), Sb0 = M' Sibar N;
  //   Uhbar += (basis cotangent of Uhat's data) + U0 Mb' (G1), Vhbar likewise (G2);
  //   Ub0 = Uhbar[:, :r] + Uhat Mb - Zbar (K1'U0) - K1 (Zbar'U0), Zbar the QR rule on
  //   Uhbar[:, r:], Kbar1 = Zbar - U0 (U0'Zbar);
  //   K reverse -> Kbar0 and the cotangents of V0's data; L likewise;
  //   Ub0 += Kbar0 S0' + (basis cotangent of U0's data, from L), Sb0 += U0'Kbar0 + Lbar0'V0.
  void quad_bug_ra_adjoint(double dt, int64_t nsub, double tol, int64_t rmax, int r1) {
    const int rr = (int)r;
    const int k2 = 2 * rr;
    const double h = dt / (double)nsub;
    const size_t d = sizeof(double);
    const int64_t mx = std::max(n, m);
    const int plen = quad_adj_part_len(rr, op_q);
    const int k3 = rr * rr * rr;
    qadj_reserve(nsub, k2);
    double* stK = qadj_rows;
    double* stL = stK + nsub * n * rr;
    double* work = stL + nsub * m * rr;  // 8 mx rr
    double* partK = work + 8 * mx * rr;
    double* partL = partK + (size_t)plen * quad_adj_grid(n, rr);
    double* swork = partL + (size_t)plen * quad_adj_grid(m, rr);
    double* kbar = work + 7 * mx * rr;
    QuadAdj* qa = qadj;
    RaAdjSmall* A = radj;
    // forward: U0's and V0's Galerkin data (kept), K and L with their sub-step starts
    quad_galerkin(V, false, rr);
    quad_galerkin(U, true, rr);
    de_gal_fresh = false;
    const size_t rsz = d * rr * rr, fsz = d * op_q * rr, tsz = d * k3;
    HIPCHECK(hipMemcpyAsync(qa->Ga0, de->Ga, rsz, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(qa->Gb0, de->Gb, rsz, hipMemcpyDeviceToDevice, stream));
    if (op_q) {
      HIPCHECK(hipMemcpyAsync(qa->gaT0, de->gaT, fsz, hipMemcpyDeviceToDevice, stream));
      HIPCHECK(hipMemcpyAsync(qa->hbT0, de->hbT, fsz, hipMemcpyDeviceToDevice, stream));
    }
    HIPCHECK(hipMemcpyAsync(qa->W0, quad_W, tsz, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(qa->T0, quad_T, tsz, hipMemcpyDeviceToDevice, stream));
    quad_init(U, n, rr, S, (int)rcap, 0, K);
    quad_rows_ex(true, rr, qa->Gb0, qa->hbT0, qa->T0, K, work, work + 2 * mx * rr, h, nsub, stK);
    quad_init(V, m, rr, S, (int)rcap, 1, L);
    quad_rows_ex(false, rr, qa->Ga0, qa->gaT0, qa->W0, L, work, work + 2 * mx * rr, h, nsub, stL);
    // the augmented bases, M, N, Sinit = M S0 N' and the completions' R factors
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      HIPCHECK(hipMemcpyAsync(ns ? Uhat : Vhat, ns ? U : V, d * rows_ * rr, hipMemcpyDeviceToDevice, stream));
      EpiArgs e;
      e.kind = ns ? EPI_RA_M : EPI_RA_N;
      if (!ns) {
        e.S = S;
        e.ld_s = (int)rcap;
      }
      double* t1 = ns ? K2 : L2;
      complete_basis(ns ? K : L, ns ? U : V, rows_, rr, t1, t1 + rows_ * rcap, (ns ? Uhat : Vhat) + rows_ * rr, e);
      HIPCHECK(hipMemcpyAsync(ns ? A->RU : A->RV, small->Rtot, rsz, hipMemcpyDeviceToDevice, stream));
    }
    quad_galerkin(Uhat, true, k2);   // Uhat'AUhat, G'Uhat, W(Uhat)
    quad_galerkin(Vhat, false, k2);  // Vhat'BVhat, H'Vhat, T(Vhat)
    quad_sstep(k2, k2, small->Sinit, k2, 0, small->Shat, k2, 1.0, h, nsub);
    ra_svd(k2, tol, rmax);
    HIPCHECK(hipStreamSynchronize(stream));
    REQUIRE(*r_host == r1, "the cotangents' rank differs from the step's new rank (recomputed: " +
                               std::to_string(*r_host) + ")");
    // Uhbar = Ub1 Us[:, :r1]' (K2 / L2) and Uhat'Ub1 (radj->Ubs / Vbs)
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      RowArgs<double> X;
      X.rows = rows_;
      X.r_out = k2;
      X.r_in = r1;
      X.X = ns ? P : Q;
      X.ldx = rows_;
      X.Y = ns ? small->Ush : small->Vsh;
      X.ldy = k2;
      X.transY = 1;
      X.out = ns ? K2 : L2;
      X.ldo = rows_;
      rows(X);
      RowArgs<double> Y;
      Y.rows = rows_;
      Y.r_out = r1;
      Y.base = ns ? P : Q;
      Y.ld_base = rows_;
      Y.reduce = 1;
      Y.W = ns ? Uhat : Vhat;
      Y.ldw = rows_;
      Y.rw = k2;
      Y.epi.kind = EPI_COPY;
      Y.epi.transpose = 1;
      Y.epi.dst = ns ? A->Ubs : A->Vbs;
      rows(Y);
    }
    // the SVD rule: the cotangent of the S sub-step's result in radj->core
    for (int stage = 0; stage < 4; ++stage)
      timed(LRI_KERNEL_SMALL, [&] { return launch_ra_adj_small(A, small, S, (int)rcap, rr, r1, stage, stream); });
    // S sub-step reverse in 2r, then G1..G4 and Sb0 = M' Sibar N from its start's cotangent
    QuadSAdjArgs SA;
    SA.k = SA.l = k2;
    SA.q = op_q;
    SA.S0 = small->Sinit;
    SA.lds0 = k2;
    SA.Sb1 = A->core;
    SA.ldsb1 = k2;
    SA.Sb0 = qa->Sib;
    SA.Ga = de->Ga;
    SA.Gb = de->Gb;
    SA.gaT = de->gaT;
    SA.hbT = de->hbT;
    SA.W = quad_W;
    SA.T = quad_T;
    SA.sgn = 1.0;
    SA.h = h;
    SA.nsub = nsub;
    SA.Gab = qa->Gab1;
    SA.Gbb = qa->Gbb1;
    SA.gaTb = qa->gaTb1;
    SA.hbTb = qa->hbTb1;
    SA.Wb = qa->Wb1;
    SA.Tb = qa->Tb1;
    SA.work = swork;
    timed(LRI_KERNEL_SMALL, [&] { return launch_quad_sadj(SA, stream); });
    HIPCHECK(hipMemcpyAsync(A->core, qa->Sib, d * k2 * k2, hipMemcpyDeviceToDevice, stream));
    timed(LRI_KERNEL_SMALL, [&] { return launch_ra_adj_small(A, small, S, (int)rcap, rr, r1, 4, stream); });
    // Uhbar (P[:, :2r]) = K2 + (basis cotangent of Uhat's data) + U0 G1; Vhbar (Q) likewise with V0 G2
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      BasisVjpArgs B = basis_vjp_args(ns, k2, ns ? Uhat : Vhat, ns ? qa->Wb1 : qa->Tb1, ns ? qa->Gab1 : qa->Gbb1,
                                      ns ? qa->gaTb1 : qa->hbTb1);
      B.base = ns ? K2 : L2;
      B.ldb = rows_;
      B.out = ns ? P : Q;
      B.ldo = rows_;
      timed(LRI_KERNEL_DE, [&] { return launch_basis_vjp(B, stream); });
      RowArgs<double> X;
      X.rows = rows_;
      X.r_out = k2;
      X.r_in = rr;
      X.base = ns ? P : Q;
      X.ld_base = rows_;
      X.X = ns ? U : V;
      X.ldx = rows_;
      X.Y = ns ? A->G1 : A->G2;
      X.ldy = rr;
      X.out = ns ? P : Q;
      X.ldo = rows_;
      rows(X);
    }
    // the completions' reverse: Kbar1 (U2) and Lbar1 (V2); Ub0 / Vb0 so far in P[:, :r] / Q[:, :r]
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      double* acc = ns ? P : Q;
      const double* Qbar = acc + rows_ * rr;
      const double* Qc = (ns ? Uhat : Vhat) + rows_ * rr;
      const double* B0 = ns ? U : V;
      const double* X1 = ns ? K : L;
      double* Zb = ns ? K2 : L2;
      double* tmp = Zb + rows_ * rcap;
      double* Xb1 = ns ? U2 : V2;
      {
        RowArgs<double> X;  // copyltu(-Qbar' Q)
        X.rows = rows_;
        X.r_out = rr;
        X.base = Qbar;
        X.ld_base = rows_;
        X.reduce = 1;
        X.W = Qc;
        X.ldw = rows_;
        X.rw = rr;
        X.epi.kind = EPI_COPYLTU;
        X.epi.dst = ns ? A->CqU : A->CqV;
        rows(X);
      }
      {
        RowArgs<double> X;  // tmp = Qbar + Q Cq
        X.rows = rows_;
        X.r_out = rr;
        X.r_in = rr;
        X.base = Qbar;
        X.ld_base = rows_;
        X.X = Qc;
        X.ldx = rows_;
        X.Y = ns ? A->CqU : A->CqV;
        X.ldy = rr;
        X.out = tmp;
        X.ldo = rows_;
        rows(X);
      }
      {
        RowArgs<double> X;  // Zbar = tmp R^{-T}; Zbar' B0
        X.rows = rows_;
        X.r_out = rr;
        X.r_in = rr;
        X.X = tmp;
        X.ldx = rows_;
        X.Y = ns ? A->RUinvT : A->RVinvT;
        X.ldy = rr;
        X.out = Zb;
        X.ldo = rows_;
        X.reduce = 1;
        X.W = B0;
        X.ldw = rows_;
        X.rw = rr;
        X.epi.kind = EPI_COPY;
        X.epi.dst = ns ? A->ZU : A->ZV;
        rows(X);
      }
      {
        RowArgs<double> X;  // K1' U0
        X.rows = rows_;
        X.r_out = rr;
        X.base = X1;
        X.ld_base = rows_;
        X.reduce = 1;
        X.W = B0;
        X.ldw = rows_;
        X.rw = rr;
        X.epi.kind = EPI_COPY;
        X.epi.dst = ns ? A->KU : A->LV;
        rows(X);
      }
      {
        RowArgs<double> X;  // Kbar1 = Zbar - U0 (Zbar'U0)'
        X.rows = rows_;
        X.r_out = rr;
        X.r_in = rr;
        X.base = Zb;
        X.ld_base = rows_;
        X.X = B0;
        X.ldx = rows_;
        X.Y = ns ? A->ZU : A->ZV;
        X.ldy = rr;
        X.transY = 1;
        X.xy_sign = -1.0;
        X.out = Xb1;
        X.ldo = rows_;
        rows(X);
      }
      // acc[:, :r] += Xhat G3 - Zbar (K1'U0) - K1 (Zbar'U0)  (V side: Vhat G4, ...)
      const double* Xs[3] = {ns ? Uhat : Vhat, Zb, X1};
      const double* Ys[3] = {ns ? A->G3 : A->G4, ns ? A->KU : A->LV, ns ? A->ZU : A->ZV};
      const int ri[3] = {k2, rr, rr};
      const int ldys[3] = {k2, rr, rr};
      const double sg[3] = {1.0, -1.0, -1.0};
      for (int u = 0; u < 3; ++u) {
        RowArgs<double> X;
        X.rows = rows_;
        X.r_out = rr;
        X.r_in = ri[u];
        X.base = acc;
        X.ld_base = rows_;
        X.X = Xs[u];
        X.ldx = rows_;
        X.Y = Ys[u];
        X.ldy = ldys[u];
        X.xy_sign = sg[u];
        X.out = acc;
        X.ldo = rows_;
        rows(X);
      }
    }
    // the K and L sub-problems' reverses: Kbar0 into K (Lbar0 into L), CK = Kbar0'U0, CL = Lbar0'V0
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      HIPCHECK(hipMemcpyAsync(kbar, ns ? U2 : V2, d * rows_ * rr, hipMemcpyDeviceToDevice, stream));
      if (ns)
        quad_rows_reverse(true, rr, qa->Gb0, qa->hbT0, qa->T0, stK, kbar, work, partK, qa->sumK, h, nsub);
      else
        quad_rows_reverse(false, rr, qa->Ga0, qa->gaT0, qa->W0, stL, kbar, work, partL, qa->sumL, h, nsub);
      RowArgs<double> D;
      D.rows = rows_;
      D.r_out = rr;
      D.base = kbar;
      D.ld_base = rows_;
      D.out = ns ? K : L;
      D.ldo = rows_;
      D.reduce = 1;
      D.W = ns ? U : V;
      D.ldw = rows_;
      D.rw = rr;
      D.epi.kind = EPI_COPY;
      D.epi.dst = ns ? adj->CK : adj->CL;
      rows(D);
    }
    // Ub0 += Kbar0 S0' + (basis cotangent of U0's data from the L reverse); Vb0 += Lbar0 S0 + ...
    for (int side = 0; side < 2; ++side) {
      const bool ns = side == 0;
      const int64_t rows_ = ns ? n : m;
      const double* sum = ns ? qa->sumL : qa->sumK;  // U0's data enter the L sub-problem, V0's the K one
      BasisVjpArgs B = basis_vjp_args(ns, rr, ns ? U : V, sum, sum + k3, sum + k3 + rr * rr);
      B.X3 = ns ? K : L;
      B.ldx3 = rows_;
      B.Y3 = S;
      B.ldy3 = (int)rcap;
      B.transY3 = ns;
      B.base = ns ? P : Q;
      B.ldb = rows_;
      B.out = ns ? P : Q;
      B.ldo = rows_;
      timed(LRI_KERNEL_DE, [&] { return launch_basis_vjp(B, stream); });
    }
    HIPCHECK(hipMemcpyAsync(adj->Sb0, A->Sb0, rsz, hipMemcpyDeviceToDevice, stream));
    timed(LRI_KERNEL_SMALL, [&] { return launch_bug_adj_post(adj, rr, stream); });  // Sb0 += CK' + CL
    timed(LRI_KERNEL_SMALL, [&] {
      return launch_quad_adj_scales(qa->sumK, qa->sumL, plen, qa->sumL + k3, qa->Ga0, qa->Gab1, de->Ga,
                                    qa->sumK + k3, qa->Gb0, qa->Gbb1, de->Gb, rr, k2, qa->scal, stream);
    });
  }

  void de_ksl_primal(double dt, int64_t nsub) {
    if (op_quad) {
      quad_ksl_primal(dt, nsub);
      return;
    }
    const int rr = (int)r;
    const double h = dt / (double)nsub;
    const EpiArgs none = de_epi();
    de_galerkin(V, false, rr, de->Gb, de->hbT);  // Bt = V0'BV0
    de_eig(de->Gb, rr, de->Wb, de->beta);
    de_substep(U, true, rr, S, (int)rcap, 0, de->Wb, de->beta, de->hbT, h, nsub, K, K2, shifted_chol(n));  // K
    cholqr_tail(K, K2, n, rr, U, false, nullptr, 0, none, 0, 3);                                            // U1, R
    de_galerkin(U, true, rr, de->Ga, de->gaT);  // At = U1'AU1
    de_eig(de->Ga, rr, de->Wa, de->alpha);
    de_sstep(rr, rr, small->Rtot, rr, 0, de->Snew, rr, -1.0, h, nsub);  // S~
    de_substep(V, false, rr, de->Snew, rr, 1, de->Wa, de->alpha, de->gaT, h, nsub, L, L2, shifted_chol(m));  // L
    cholqr_tail(L, L2, m, rr, V, true, nullptr, 0, none, 2, 3);  // V1, S = R'
  }

  // dual_LT_step! with the DE sub-problems (projector_splitting.jl:52-80,166-189)
  void de_ksl_dual(double dt, int64_t nsub) {
    if (op_quad) {
      quad_ksl_dual(dt, nsub);
      return;
    }
    const int rr = (int)r;
    const double h = dt / (double)nsub;
    const EpiArgs none = de_epi();
    de_galerkin(U, true, rr, de->Ga, de->gaT);  // At = U0'AU0
    de_eig(de->Ga, rr, de->Wa, de->alpha);
    de_substep(V, false, rr, S, (int)rcap, 1, de->Wa, de->alpha, de->gaT, h, nsub, L, L2, shifted_chol(m));  // L
    cholqr_tail(L, L2, m, rr, V, true, nullptr, 0, none, 0, 3);                                              // V1, R
    de_galerkin(V, false, rr, de->Gb, de->hbT);  // Bt = V1'BV1
    de_eig(de->Gb, rr, de->Wb, de->beta);
    de_sstep(rr, rr, small->Rtot, rr, 1, de->Snew, rr, -1.0, h, nsub);  // S~ from R'
    de_substep(U, true, rr, de->Snew, rr, 0, de->Wb, de->beta, de->hbT, h, nsub, K, K2, shifted_chol(n));  // K
    cholqr_tail(K, K2, n, rr, U, false, nullptr, 0, none, 1, 3);  // U1, S = R
  }

  // unconventional_step! with the DE sub-problems (unconventional.jl:51-79,133-157)
  // Scratch columns of the L sub-problem of a step that runs K and L together:
  // right after the K ones in K2 when both sides share one eigenbasis (so that
  // de_run transforms both at once), else L2.
  double* de_lscratch(int k) {
    const bool fuse = n == m && spec[0] && spec[1] && specQ[0] == specQ[1] && 2 * k <= 64;
    return fuse ? K2 + n * k : L2;
  }

  void de_bug(double dt, int64_t nsub) {
    if (op_quad) {
      quad_bug(dt, nsub);
      return;
    }
    const int rr = (int)r;
    const double h = dt / (double)nsub;
    if (!de_gal_fresh) {
      de_galerkin(V, false, rr, de->Gb, de->hbT);  // Bt = V0'BV0
      de_galerkin(U, true, rr, de->Ga, de->gaT);   // At = U0'AU0
      de_eig2(rr, rr);
    }
    de_gal_fresh = false;
    // the K and L sub-problems are independent: one RK4 launch for both
    double* lscr = de_lscratch(rr);
    const RK4ColArgs Pk = de_prepare(U, true, rr, S, (int)rcap, 0, de->Wb, de->beta, de->hbT, 0, h, nsub, K2);
    const RK4ColArgs Pl = de_prepare(V, false, rr, S, (int)rcap, 1, de->Wa, de->alpha, de->gaT, 1, h, nsub, lscr);
    de_run(Pk, Pl);
    de_finish(true, rr, de->Wb, K, K2, shifted_chol(n));
    EpiArgs eM = de_epi(EPI_BUG_M);  // M = U1' U0
    cholqr_tail(K, K2, n, rr, U2, false, U, rr, eM, 0, 3);
    de_finish(false, rr, de->Wa, L, lscr, shifted_chol(m));
    EpiArgs eN = de_epi(EPI_BUG_N);  // N = V1' V0, Sinit = M S0 N'
    eN.S = S;
    eN.ld_s = (int)rcap;
    cholqr_tail(L, L2, m, rr, V2, true, V, rr, eN, 0, 3);
    de_galerkin(U2, true, rr, de->Ga, de->gaT);   // At = U1'AU1
    de_galerkin(V2, false, rr, de->Gb, de->hbT);  // Bt = V1'BV1
    de_eig2(rr, rr);
    de_sstep(rr, rr, small->Sinit, rr, 0, S, (int)rcap, 1.0, h, nsub);
    std::swap(U, U2);
    std::swap(V, V2);
    de_gal_fresh = true;  // Ga, Gb, gaT, hbT and their eigenpairs now belong to U, V
  }

  // rankadaptive_unconventional_step! with the DE sub-problems
  // (rank_adaptive_unconventional.jl:59-86,194-233)
  // complete = false: stop after queueing the SVD (ra_rotate finishes the step)
  int64_t de_bug_ra(double dt, int64_t nsub, double tol, int64_t rmax, bool complete = true) {
    if (op_quad) return quad_bug_ra(dt, nsub, tol, rmax, complete);
    const int rr = (int)r;
    const int k2 = 2 * rr;
    const double h = dt / (double)nsub;
    EpiArgs none;
    de_galerkin(V, false, rr, de->Gb, de->hbT);  // Bt = V0'BV0
    de_galerkin(U, true, rr, de->Ga, de->gaT);   // At = U0'AU0
    de_eig2(rr, rr);
    double* lscr = de_lscratch(rr);
    const RK4ColArgs Pk = de_prepare(U, true, rr, S, (int)rcap, 0, de->Wb, de->beta, de->hbT, 0, h, nsub, K2);
    const RK4ColArgs Pl = de_prepare(V, false, rr, S, (int)rcap, 1, de->Wa, de->alpha, de->gaT, 1, h, nsub, lscr);
    de_run(Pk, Pl);
    de_finish(true, rr, de->Wb, K, K2, none);
    de_finish(false, rr, de->Wa, L, lscr, none);
    dcopy(Uhat, U, n * rr);
    EpiArgs eM;
    eM.kind = EPI_RA_M;
    complete_basis(K, U, n, rr, K2, K2 + n * rcap, Uhat + n * rr, eM);
    dcopy(Vhat, V, m * rr);
    EpiArgs eN;
    eN.kind = EPI_RA_N;  // Sinit = M S0 N' (2r x 2r)
    eN.S = S;
    eN.ld_s = (int)rcap;
    complete_basis(L, V, m, rr, L2, L2 + m * rcap, Vhat + m * rr, eN);
    de_galerkin(Uhat, true, k2, de->Ga, de->gaT);
    de_galerkin(Vhat, false, k2, de->Gb, de->hbT);
    de_eig2(k2, k2);
    de_sstep(k2, k2, small->Sinit, k2, 0, small->Shat, k2, 1.0, h, nsub);
    ra_svd(k2, tol, rmax);
    return complete ? ra_rotate() : 0;
  }

  // Constant-coefficient sides get their analytic eigenbasis (used by de_run);
  // G and H are projected once: specF = Q' G (or Q' H).
  void setup_spectral(const double* ad, const double* ae, bool aper, const double* bd, const double* be, bool bper) {
    auto constant = [](const double* d, const double* e, int64_t len, bool per, double* d0, double* e0) {
      *d0 = d[0];
      *e0 = (len > 1 || per) ? e[0] : 0.0;
      for (int64_t i = 1; i < len; ++i)
        if (d[i] != *d0) return false;
      const int64_t ne = per ? len : len - 1;
      for (int64_t i = 1; i < ne; ++i)
        if (e[i] != *e0) return false;
      return true;
    };
    HIPCHECK(hipStreamSynchronize(stream));
    for (int side = 0; side < 2; ++side) {
      if (side == 1 && spec_alias) specQ[1] = specLam[1] = nullptr;  // owned by side 0
      for (double** p : {&specQ[side], &specLam[side], &specF[side]})
        if (*p) {
          HIPCHECK(hipFree(*p));
          *p = nullptr;
        }
      spec[side] = false;
    }
    spec_alias = false;
    double da, ea, db, eb;
    const bool ca = constant(ad, ae, n, aper, &da, &ea);
    const bool cb = constant(bd, be, m, bper, &db, &eb);
    for (int side = 0; side < 2; ++side) {
      const bool c = side == 0 ? ca : cb;
      if (!c) continue;
      const int64_t rn = side == 0 ? n : m;
      if (side == 1 && ca && m == n && aper == bper && da == db && ea == eb) {
        spec_alias = true;
        specQ[1] = specQ[0];
        specLam[1] = specLam[0];
        HIPCHECK(hipMalloc((void**)&specF[1], sizeof(double) * rn * LRI_MAX_FORCE));
      } else {
        HIPCHECK(hipMalloc((void**)&specQ[side], sizeof(double) * rn * rn));
        HIPCHECK(hipMalloc((void**)&specLam[side], sizeof(double) * rn));
        HIPCHECK(hipMalloc((void**)&specF[side], sizeof(double) * rn * LRI_MAX_FORCE));
        HIPCHECK(launch_eig_tridiag_const(specQ[side], specLam[side], rn, side == 0 ? da : db, side == 0 ? ea : eb,
                                          side == 0 ? aper : bper, stream));
      }
      HIPCHECK(hipMemsetAsync(specF[side], 0, sizeof(double) * rn * LRI_MAX_FORCE, stream));
      if (op_q > 0) spectral_gemm(true, specQ[side], rn, side == 0 ? op_G : op_H, op_q, specF[side]);
      spec[side] = true;
    }
    HIPCHECK(hipStreamSynchronize(stream));
  }

  // X (rows x k, ld rows) <- op(Q) X for a side's analytic eigenbasis, through tmp
  void basis_transform(int side, bool trans, double* X, int k, double* tmp) {
    const int64_t rn = side == 0 ? n : m;
    spectral_gemm(trans, specQ[side], rn, X, k, tmp);
    HIPCHECK(hipMemcpyAsync(X, tmp, sizeof(double) * rn * k, hipMemcpyDeviceToDevice, stream));
  }
  void enter_resident() {
    if (resident) return;
    drop_graphs();
    REQUIRE(op_set && spec[0] && spec[1] && !op_quad,
            "the resident eigenbasis needs a linear operator with constant coefficients on both sides");
    if (!res_zero) {
      HIPCHECK(hipMalloc((void**)&res_zero, sizeof(double) * std::max(n, m)));
      HIPCHECK(hipMemset(res_zero, 0, sizeof(double) * std::max(n, m)));
    }
    double** ops[6] = {&op_ad, &op_ae, &op_bd, &op_be, &op_G, &op_H};
    for (int i = 0; i < 6; ++i) res_saved[i] = *ops[i];
    res_saved_per[0] = op_aper;
    res_saved_per[1] = op_bper;
    op_ad = specLam[0];
    op_ae = res_zero;
    op_bd = specLam[1];
    op_be = res_zero;
    op_G = specF[0];
    op_H = specF[1];
    op_aper = op_bper = 0;
    if (r >= 1) {
      basis_transform(0, true, U, (int)r, K2);
      basis_transform(1, true, V, (int)r, L2);
    }
    resident = true;
    pkN_mode = pkT_mode = -1;
  }
  void leave_resident() {
    if (!resident) return;
    drop_graphs();
    double** ops[6] = {&op_ad, &op_ae, &op_bd, &op_be, &op_G, &op_H};
    for (int i = 0; i < 6; ++i) *ops[i] = res_saved[i];
    op_aper = res_saved_per[0];
    op_bper = res_saved_per[1];
    resident = false;
    if (r >= 1) {
      basis_transform(0, false, U, (int)r, K2);
      basis_transform(1, false, V, (int)r, L2);
    }
    pkN_mode = pkT_mode = -1;
  }
  // U = Qa U^, V = Qb V^ of the resident state (r columns) into K2 and L2
  void resident_out(const double* Uh, const double* Vh, int rr) {
    spectral_gemm(false, specQ[0], n, Uh, rr, K2);
    spectral_gemm(false, specQ[1], m, Vh, rr, L2);
  }

  void check_de(double dt, int64_t nsub) {
    REQUIRE(r >= 1, "state not set (call lri_set_state)");
    REQUIRE(op_set, "no operator (call lri_set_operator_lyapunov)");
    REQUIRE(dt > 0.0 && std::isfinite(dt), "dt must be positive and finite");
    REQUIRE(nsub >= 1, "substeps must be >= 1");
    REQUIRE(!op_quad || r <= kQuadMaxRank, "the quadratic operator needs rank <= LRI_QUAD_MAX_RANK");
  }

  int read_flags() {
    int f = 0;
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipMemcpy(&f, flags, sizeof(int), hipMemcpyDeviceToHost));
    if (f) HIPCHECK(hipMemset(flags, 0, sizeof(int)));
    if (timing) drain_timing();
    if (f & FLAG_NONFINITE) return LRI_ENONFINITE;
    if (f & FLAG_BREAKDOWN) return LRI_EBREAKDOWN;
    return LRI_OK;
  }
};

namespace {

template <bool KeepDeGalerkin = false, class F>
int guarded(lri_ctx* ctx, F&& f) {
  if (!ctx) return LRI_EINVAL;
  if (!KeepDeGalerkin) ctx->de_gal_fresh = false;
  try {
    HIPCHECK(hipSetDevice(ctx->device));
    int st = f();
    if (st != LRI_OK && ctx->err.empty()) ctx->err = lri_status_string(st);
    return st;
  } catch (const Failure& e) {
    ctx->err = e.msg;
    return e.status;
  } catch (const std::bad_alloc&) {
    ctx->err = "host allocation failed";
    return LRI_EOOM;
  } catch (...) {
    ctx->err = "unexpected C++ exception";
    return LRI_EHIP;
  }
}

size_t dtype_size(int dt) { return dt == LRI_F64 ? 8 : (dt == LRI_F32 ? 4 : 2); }

void free_ctx(lri_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->resident) {  // the operator's own arrays (the resident ones alias specLam / specF)
    double** ops[6] = {&c->op_ad, &c->op_ae, &c->op_bd, &c->op_be, &c->op_G, &c->op_H};
    for (int i = 0; i < 6; ++i) *ops[i] = c->res_saved[i];
    c->resident = false;
  }
  if (c->res_zero) (void)hipFree(c->res_zero);
  for (void* p : c->owned) (void)hipFree(p);
  void* bufs[] = {c->U,      c->V,      c->S,      c->U2,      c->V2,      c->K,      c->K2,     c->P,
                  c->L,      c->L2,     c->Q,      c->Uhat,    c->Vhat,    c->pkN_f64, c->pkT_f64, c->pkN_f32,
                  c->pkT_f32, c->pkN_hi, c->pkN_lo, c->pkT_hi,  c->pkT_lo,  c->partN,  c->partT,  c->gpart,
                  c->flags,  c->small,  c->svd_work, c->op_ad,  c->op_ae,   c->op_bd,   c->op_be,  c->op_G,
                  c->op_H,   c->de,     c->specQ[0], c->specLam[0], c->specF[0], c->op_dl, c->op_dd, c->op_du,
                  c->quad_part, c->quad_T, c->quad_W, c->quad_work};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (c->specF[1]) (void)hipFree(c->specF[1]);
  if (!c->spec_alias)
    for (void* p : {(void*)c->specQ[1], (void*)c->specLam[1]})
      if (p) (void)hipFree(p);
  if (c->r_host) (void)hipHostFree(c->r_host);
  if (c->adj) (void)hipFree(c->adj);
  if (c->radj) (void)hipFree(c->radj);
  if (c->qadj) (void)hipFree(c->qadj);
  if (c->qadj_rows) (void)hipFree(c->qadj_rows);
  if (c->traj) (void)hipFree(c->traj);
  if (c->backup) (void)hipFree(c->backup);
  for (auto& g : c->ra_graphs) (void)hipGraphExecDestroy(g.exec);
  for (auto& t : c->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

}  // namespace

// ---------------------------------------------------------------- batches
namespace {

// Installs a recorder on the calling thread for its lifetime.
struct RecorderScope {
  Recorder* prev;
  explicit RecorderScope(Recorder* r) : prev(current_recorder()) { current_recorder() = r; }
  ~RecorderScope() { current_recorder() = prev; }
};

// One phase of a merged batch step.  Each member's launches are recorded once
// per key (the step's parameters, the member's rank and its epoch) and their
// argument blocks kept at fixed slots of a device pool: member i owns slots
// i * cap .. (i + 1) * cap - 1, launch p of its sequence slot i * cap + p.  The
// merged list runs position p of every member's sequence as one launch per
// (kernel, variant, threads), with the largest LDS size of its members; the
// launches of one position touch disjoint members' buffers, and positions run
// in order on one stream, so each member sees its own launches in its own order.
struct MergedPhase {
  using Key = std::array<double, 8>;
  struct Launch {
    int kid, variant;
    unsigned block, total;
    size_t lds;
    int count;
    size_t meta_off;  // its workgroup offsets, slots and workgroup members in `meta` (2 count + 1 + total ints)
  };
  int count = 0;
  int cap = 0;
  std::vector<std::vector<LaunchRec>> recs;
  std::vector<Key> keys;
  std::vector<char> have, dirty, part, plan_part;
  unsigned char* pool = nullptr;       // device, count * cap argument blocks
  unsigned char* pool_host = nullptr;  // pinned staging of the same layout
  std::vector<Launch> launches;
  std::vector<int> meta_vec;
  int* meta = nullptr;       // device
  int* meta_host = nullptr;  // pinned
  size_t meta_cap = 0;       // ints
  bool plan_stale = true;
  uint64_t plan_gen = 0, seen_gen = ~0ull, exec_gen = ~0ull;
  hipGraphExec_t exec = nullptr;

  void init(int n) {
    count = n;
    recs.assign(n, {});
    keys.assign(n, Key{});
    have.assign(n, 0);
    dirty.assign(n, 0);
    part.assign(n, 0);
  }
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (pool) (void)hipFree(pool);
    if (pool_host) (void)hipHostFree(pool_host);
    if (meta) (void)hipFree(meta);
    if (meta_host) (void)hipHostFree(meta_host);
    exec = nullptr;
    pool = pool_host = nullptr;
    meta = meta_host = nullptr;
    cap = 0;
    meta_cap = 0;
  }
  // Member i's launches for `key`: recorded by f() on a miss, else the cached ones.
  template <class F>
  void record(int i, const Key& key, F&& f) {
    if (have[i] && keys[i] == key) return;
    have[i] = 0;
    Recorder R;
    {
      RecorderScope scope(&R);
      f();
    }
    recs[i] = std::move(R.recs);
    keys[i] = key;
    have[i] = 1;
    dirty[i] = 1;
    plan_stale = true;
  }
  // The merged list of the participating members; plan_gen changes only when
  // the list itself does (new argument blocks at the same slots leave it).
  void build_plan() {
    std::vector<Launch> old_launches;
    std::vector<int> old_meta;
    old_launches.swap(launches);
    old_meta.swap(meta_vec);
    size_t maxlen = 0;
    for (int i = 0; i < count; ++i)
      if (part[i]) maxlen = std::max(maxlen, recs[i].size());
    struct Group {
      int kid, variant;
      unsigned block;
      size_t lds;
      std::vector<int> members;
    };
    std::vector<Group> groups;
    for (size_t p = 0; p < maxlen; ++p) {
      groups.clear();
      for (int i = 0; i < count; ++i) {
        if (!part[i] || p >= recs[i].size() || recs[i][p].grid == 0) continue;
        const LaunchRec& L = recs[i][p];
        Group* g = nullptr;
        for (Group& x : groups)
          if (x.kid == L.kid && x.variant == L.variant && x.block == L.block) g = &x;
        if (!g) {
          groups.push_back(Group{L.kid, L.variant, L.block, 0, {}});
          g = &groups.back();
        }
        g->lds = std::max<size_t>(g->lds, L.lds);
        g->members.push_back(i);
      }
      for (const Group& g : groups) {
        Launch ln{g.kid, g.variant, g.block, 0, g.lds, (int)g.members.size(), meta_vec.size()};
        unsigned acc = 0;
        for (int i : g.members) {
          meta_vec.push_back((int)acc);
          acc += recs[i][p].grid;
        }
        meta_vec.push_back((int)acc);
        for (int i : g.members) meta_vec.push_back(i * cap + (int)p);
        for (size_t m = 0; m < g.members.size(); ++m)  // the member of every workgroup
          meta_vec.insert(meta_vec.end(), recs[g.members[m]][p].grid, (int)m);
        ln.total = acc;
        launches.push_back(ln);
      }
    }
    plan_part = part;
    plan_stale = false;
    bool same = old_meta == meta_vec && old_launches.size() == launches.size();
    for (size_t k = 0; same && k < launches.size(); ++k) {
      const Launch &a = launches[k], &b = old_launches[k];
      same = a.kid == b.kid && a.variant == b.variant && a.block == b.block && a.total == b.total && a.lds == b.lds &&
             a.count == b.count && a.meta_off == b.meta_off;
    }
    if (!same) ++plan_gen;
    return;
  }
  // Stage the changed members' argument blocks and, after a change, the merged
  // list.  The pinned staging is rewritten only after the stream has passed the
  // copies that read it (a batch step waits on its stream between phases).
  void upload(hipStream_t st) {
    size_t need = 1;
    for (int i = 0; i < count; ++i)
      if (part[i]) need = std::max(need, recs[i].size());
    if ((int)need > cap) {
      HIPCHECK(hipStreamSynchronize(st));
      if (pool) HIPCHECK(hipFree(pool));
      if (pool_host) HIPCHECK(hipHostFree(pool_host));
      pool = pool_host = nullptr;
      cap = (int)((need + need / 4 + 15) / 16 * 16);
      const size_t bytes = (size_t)count * cap * kRecArgBytes;
      HIPCHECK(hipMalloc((void**)&pool, bytes));
      HIPCHECK(hipHostMalloc((void**)&pool_host, bytes, hipHostMallocDefault));
      for (int i = 0; i < count; ++i) dirty[i] = have[i];
      plan_stale = true;
      ++plan_gen;  // the launches' pool pointer changed
    }
    int lo = count, hi = -1, nd = 0;
    for (int i = 0; i < count; ++i) {
      if (!dirty[i]) continue;
      for (size_t p = 0; p < recs[i].size(); ++p)
        std::memcpy(pool_host + ((size_t)i * cap + p) * kRecArgBytes, recs[i][p].args, kRecArgBytes);
      lo = std::min(lo, i);
      hi = i;
      ++nd;
    }
    if (nd > 0 && 4 * nd >= hi - lo + 1) {  // one copy over the changed span
      const size_t off = (size_t)lo * cap * kRecArgBytes, bytes = (size_t)(hi - lo + 1) * cap * kRecArgBytes;
      HIPCHECK(hipMemcpyAsync(pool + off, pool_host + off, bytes, hipMemcpyHostToDevice, st));
    } else {
      for (int i = lo; i <= hi; ++i)
        if (dirty[i]) {
          const size_t off = (size_t)i * cap * kRecArgBytes;
          HIPCHECK(hipMemcpyAsync(pool + off, pool_host + off, recs[i].size() * kRecArgBytes, hipMemcpyHostToDevice,
                                  st));
        }
    }
    std::fill(dirty.begin(), dirty.end(), 0);
    if (plan_stale || part != plan_part) {
      const uint64_t gen = plan_gen;
      build_plan();
      if (gen == plan_gen && meta) return;  // the same list: the device copy stands
      if (meta_vec.size() > meta_cap) {
        HIPCHECK(hipStreamSynchronize(st));
        if (meta) HIPCHECK(hipFree(meta));
        if (meta_host) HIPCHECK(hipHostFree(meta_host));
        meta = meta_host = nullptr;
        meta_cap = meta_vec.size() + meta_vec.size() / 4 + 64;
        HIPCHECK(hipMalloc((void**)&meta, sizeof(int) * meta_cap));
        HIPCHECK(hipHostMalloc((void**)&meta_host, sizeof(int) * meta_cap, hipHostMallocDefault));
        ++plan_gen;  // the launches' pointers changed
      }
      if (!meta_vec.empty()) {
        std::memcpy(meta_host, meta_vec.data(), sizeof(int) * meta_vec.size());
        HIPCHECK(hipMemcpyAsync(meta, meta_host, sizeof(int) * meta_vec.size(), hipMemcpyHostToDevice, st));
      }
    }
  }
  void dispatch(const Launch& ln, hipStream_t st) const {
    const int* start = meta + ln.meta_off;
    const dim3 g(ln.total), b(ln.block);
    switch (ln.kid) {
      case KID_SET_DIAG:
        hipLaunchKernelGGL(k_set_diag_batch, g, b, ln.lds, st, pool, start, ln.count);
        HIPCHECK(hipGetLastError());
        break;
      case KID_COPY:
        hipLaunchKernelGGL(k_copy_batch, g, b, ln.lds, st, pool, start, ln.count);
        HIPCHECK(hipGetLastError());
        break;
      case KID_ROWS_D:
      case KID_ROWS_F:
      case KID_REDUCE:
      case KID_EPILOGUE:
      case KID_JACOBI_WAVE:
      case KID_JACOBI_WG:
        HIPCHECK(launch_rowops_batched(ln.kid, ln.variant, pool, start, ln.count, ln.total, ln.block, ln.lds, st));
        break;
      default:
        HIPCHECK(launch_de_batched(ln.kid, ln.variant, pool, start, ln.count, ln.total, ln.block, ln.lds, st));
    }
  }
  // Run the merged list: directly the first time a list is seen, then as a
  // captured graph while it stays the same (the graph reads the pool and the
  // offsets at replay, so changed argument blocks need no new capture).
  void run(hipStream_t st, bool graphs) {
    if (graphs && exec && exec_gen == plan_gen) {
      HIPCHECK(hipGraphLaunch(exec, st));
      return;
    }
    if (graphs && seen_gen == plan_gen && !launches.empty()) {
      if (exec) (void)hipGraphExecDestroy(exec);
      exec = nullptr;
      exec_gen = ~0ull;
      hipGraph_t graph = nullptr;
      HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        for (const Launch& ln : launches) dispatch(ln, st);
      } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      HIPCHECK(hipStreamEndCapture(st, &graph));
      hipGraphExec_t x = nullptr;
      const hipError_t e = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHECK(e);
      exec = x;
      exec_gen = plan_gen;
      HIPCHECK(hipGraphLaunch(exec, st));
      return;
    }
    seen_gen = plan_gen;
    for (const Launch& ln : launches) dispatch(ln, st);
  }
};

}  // namespace

// Independent problems of uneven n, m and rank, one context each.  With merged
// launches (the default) the members share one stream and a rank-adaptive
// batch step runs as one launch per stage over all members (MergedPhase);
// without, each member keeps its own stream and a batch step queues every
// member's step before it waits for any of them.
struct lri_batch {
  int device = 0;
  std::vector<lri_ctx*> ctx;
  std::string err;
  bool graphs = true;  // captured graphs (merged lists, or each member's step)
  std::vector<int> status;  // each member's status from the last batch call
  bool merged = false;
  hipStream_t stream = nullptr;  // shared by the members while merged
  MergedPhase ph1, ph2;         // up to the SVD; the rotation to the new rank
  MemberWords* words = nullptr;  // device: each member's rank and flag words
  int* word_out = nullptr;       // device: 2 count gathered words
  int* word_host = nullptr;      // pinned
  int64_t merged_steps = 0;      // batch steps run as merged launches
};

namespace {
// run f(i, ctx) on every member whose status is still OK (all members when
// `fresh`), keep each member's status, and report the first failure as
// "problem i: ..." (a failing member does not stop the others)
template <class F>
int batch_each(lri_batch* b, F&& f, bool fresh = true) {
  if (!b) return LRI_EINVAL;
  if (fresh) b->status.assign(b->ctx.size(), LRI_OK);
  int first = LRI_OK;
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    if (b->status[i] != LRI_OK) continue;
    const int st = f((int64_t)i, b->ctx[i]);
    b->status[i] = st;
    if (st != LRI_OK && first == LRI_OK) {
      b->err = "problem " + std::to_string(i) + ": " + lri_last_error(b->ctx[i]);
      first = st;
    }
  }
  return first;
}

// Point every member at the batch's shared stream (on) or back at its own
// (off), each stream drained first.
void batch_set_merged(lri_batch* b, bool on) {
  if (on == b->merged) return;
  HIPCHECK(hipSetDevice(b->device));
  for (lri_ctx* c : b->ctx) HIPCHECK(hipStreamSynchronize(c->stream));
  if (on) {
    const int count = (int)b->ctx.size();
    if (!b->stream) HIPCHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    if (!b->words) {
      std::vector<MemberWords> w(count);
      for (int i = 0; i < count; ++i) w[i] = MemberWords{&b->ctx[i]->small->r_new, b->ctx[i]->flags};
      HIPCHECK(hipMalloc((void**)&b->words, sizeof(MemberWords) * count));
      HIPCHECK(hipMemcpy(b->words, w.data(), sizeof(MemberWords) * count, hipMemcpyHostToDevice));
      HIPCHECK(hipMalloc((void**)&b->word_out, sizeof(int) * 2 * count));
      HIPCHECK(hipHostMalloc((void**)&b->word_host, sizeof(int) * 2 * count, hipHostMallocDefault));
      b->ph1.init(count);
      b->ph2.init(count);
    }
    for (lri_ctx* c : b->ctx) c->stream = b->stream;
  } else {
    for (lri_ctx* c : b->ctx) c->stream = c->own_stream;
  }
  b->merged = on;
}

void batch_free(lri_batch* b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (lri_ctx* c : b->ctx) {
    c->stream = c->own_stream;
    free_ctx(c);
  }
  b->ph1.release();
  b->ph2.release();
  if (b->words) (void)hipFree(b->words);
  if (b->word_out) (void)hipFree(b->word_out);
  if (b->word_host) (void)hipHostFree(b->word_host);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // namespace

extern "C" {

int lri_version(void) { return 100; }

const char* lri_status_string(int s) {
  switch (s) {
    case LRI_OK: return "ok";
    case LRI_EINVAL: return "invalid argument";
    case LRI_EHIP: return "HIP runtime error";
    case LRI_EOOM: return "out of device memory";
    case LRI_EBREAKDOWN: return "numerically singular Gram matrix in CholeskyQR2";
    case LRI_ENONFINITE: return "non-finite values reached a Gram matrix";
    case LRI_ENOTSUP: return "not supported";
    default: return "unknown status";
  }
}

const char* lri_last_error(const lri_ctx* ctx) {
  if (!ctx) return g_create_error.c_str();
  return ctx->err.c_str();
}

namespace {
bool shape_ok(int64_t n, int64_t m, int64_t rcap) {
  return n >= 1 && m >= 1 && rcap >= 1 && rcap <= LRI_MAX_RANK && rcap <= n && rcap <= m;
}
}  // namespace

int lri_create(lri_ctx** out, int device, int64_t n, int64_t m, int64_t rcap) {
  g_create_error.clear();
  if (!out) {
    g_create_error = "out is NULL";
    return LRI_EINVAL;
  }
  *out = nullptr;
  if (!shape_ok(n, m, rcap)) {
    g_create_error = "need n, m >= 1 and 1 <= rcap <= min(n, m, LRI_MAX_RANK)";
    return LRI_EINVAL;
  }
  lri_ctx* c = nullptr;
  try {
    c = new lri_ctx();
    c->device = device;
    c->n = n;
    c->m = m;
    c->rcap = rcap;
    HIPCHECK(hipSetDevice(device));
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = c->stream;
    auto alloc = [&](void** p, size_t bytes) {
      HIPCHECK(hipMalloc(p, bytes));
      HIPCHECK(hipMemset(*p, 0, bytes));
    };
    const size_t d = sizeof(double);
    const int64_t r2 = 2 * rcap;
    alloc((void**)&c->U, d * n * rcap);
    alloc((void**)&c->V, d * m * rcap);
    alloc((void**)&c->S, d * rcap * rcap);
    alloc((void**)&c->U2, d * n * rcap);
    alloc((void**)&c->V2, d * m * rcap);
    alloc((void**)&c->K, d * n * r2);
    alloc((void**)&c->K2, d * n * r2);
    alloc((void**)&c->P, d * n * r2);
    alloc((void**)&c->L, d * m * r2);
    alloc((void**)&c->L2, d * m * r2);
    alloc((void**)&c->Q, d * m * r2);
    alloc((void**)&c->Uhat, d * n * r2);
    alloc((void**)&c->Vhat, d * m * r2);
    c->ldm_pk = pad_to(m, 64);
    c->ldn_pk = pad_to(n, 64);
    const int64_t rbmax = 128;
    alloc((void**)&c->pkN_f64, d * m * rbmax);
    alloc((void**)&c->pkT_f64, d * n * rbmax);
    alloc((void**)&c->pkN_f32, 4 * c->ldm_pk * rbmax);
    alloc((void**)&c->pkT_f32, 4 * c->ldn_pk * rbmax);  // pass-T order: whole 32-row blocks
    alloc((void**)&c->pkN_hi, 2 * c->ldm_pk * rbmax);
    alloc((void**)&c->pkN_lo, 2 * c->ldm_pk * rbmax);
    alloc((void**)&c->pkT_hi, 2 * c->ldn_pk * rbmax);
    alloc((void**)&c->pkT_lo, 2 * c->ldn_pk * rbmax);
    // partial buffers: the largest plan over modes, directions and widths
    size_t pb = 0;
    for (int mode = 0; mode < 3; ++mode)
      for (int rb : {16, 32, 48, 64, 128}) {
        if (mode != PASS_F64 && (rb == 16 || rb == 48)) continue;
        int s;
        int64_t ch;
        size_t es = mode == PASS_F64 ? 8 : 4;
        pb = std::max(pb, es * (size_t)plan_pass(mode, false, n, m, rb, &s, &ch));
        pb = std::max(pb, es * (size_t)plan_pass(mode, true, m, n, rb, &s, &ch));
        if (mode == PASS_F64 && rb <= 64)  // the matrix-ODE eigenbasis transforms (n x n, m x m)
          for (int64_t d : {n, m})
            for (bool tr : {false, true}) pb = std::max(pb, es * (size_t)plan_pass(mode, tr, d, d, rb, &s, &ch));
      }
    c->part_bytes = pb;
    alloc(&c->partN, pb);
    alloc(&c->partT, pb);
    alloc((void**)&c->gpart, d * 256 * kMaxAug * kMaxAug);
    alloc((void**)&c->flags, 64);
    alloc((void**)&c->small, sizeof(SmallState));
    alloc((void**)&c->svd_work, d * kSvdWorkDoubles);
    HIPCHECK(hipHostMalloc((void**)&c->r_host, sizeof(int), hipHostMallocDefault));
    *c->r_host = 0;
    HIPCHECK(hipDeviceSynchronize());
  } catch (const Failure& e) {
    g_create_error = e.msg;
    free_ctx(c);
    return e.status;
  } catch (const std::bad_alloc&) {
    g_create_error = "host allocation failed";
    free_ctx(c);
    return LRI_EOOM;
  }
  *out = c;
  return LRI_OK;
}

int lri_destroy(lri_ctx* ctx) {
  free_ctx(ctx);
  return LRI_OK;
}

int lri_set_precision(lri_ctx* ctx, int precision) {
  return guarded<true>(ctx, [&] {
    REQUIRE(precision == LRI_PREC_F64 || precision == LRI_PREC_MIXED, "precision must be LRI_PREC_F64 or LRI_PREC_MIXED");
    ctx->prec = precision;
    return LRI_OK;
  });
}

int lri_set_state(lri_ctx* ctx, int64_t r, const double* U, int64_t ldu, const double* S, int64_t lds,
                  const double* V, int64_t ldv) {
  return guarded(ctx, [&] {
    REQUIRE(r >= 1 && r <= ctx->rcap, "rank must be in [1, rcap]");
    REQUIRE(U && S && V, "U, S and V are required");
    REQUIRE(ldu >= ctx->n && ldv >= ctx->m && lds >= r, "leading dimension too small");
    const size_t d = sizeof(double);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy2D(ctx->U, d * ctx->n, U, d * ldu, d * ctx->n, r, hipMemcpyDefault));
    HIPCHECK(hipMemcpy2D(ctx->V, d * ctx->m, V, d * ldv, d * ctx->m, r, hipMemcpyDefault));
    HIPCHECK(hipMemset(ctx->S, 0, d * ctx->rcap * ctx->rcap));
    HIPCHECK(hipMemcpy2D(ctx->S, d * ctx->rcap, S, d * lds, d * r, r, hipMemcpyDefault));
    ctx->r = r;
    ctx->pkN_mode = ctx->pkT_mode = -1;
    if (ctx->resident) {  // into the eigenbasis
      ctx->basis_transform(0, true, ctx->U, (int)r, ctx->K2);
      ctx->basis_transform(1, true, ctx->V, (int)r, ctx->L2);
    }
    return LRI_OK;
  });
}

int lri_get_state(lri_ctx* ctx, int64_t* r_out, double* U, int64_t ldu, double* S, int64_t lds, double* V,
                  int64_t ldv) {
  return guarded<true>(ctx, [&] {
    const int64_t r = ctx->r;
    const double* Us = ctx->U;
    const double* Vs = ctx->V;
    if (ctx->resident && r >= 1 && (U || V)) {  // out of the eigenbasis
      ctx->resident_out(ctx->U, ctx->V, (int)r);
      Us = ctx->K2;
      Vs = ctx->L2;
    }
    int st = ctx->read_flags();
    const size_t d = sizeof(double);
    if (r_out) *r_out = r;
    if (r >= 1) {
      if (U) {
        REQUIRE(ldu >= ctx->n, "ldu too small");
        HIPCHECK(hipMemcpy2D(U, d * ldu, Us, d * ctx->n, d * ctx->n, r, hipMemcpyDefault));
      }
      if (V) {
        REQUIRE(ldv >= ctx->m, "ldv too small");
        HIPCHECK(hipMemcpy2D(V, d * ldv, Vs, d * ctx->m, d * ctx->m, r, hipMemcpyDefault));
      }
      if (S) {
        REQUIRE(lds >= r, "lds too small");
        HIPCHECK(hipMemcpy2D(S, d * lds, ctx->S, d * ctx->rcap, d * r, r, hipMemcpyDefault));
      }
    }
    return st;
  });
}

int lri_get_rank(lri_ctx* ctx, int64_t* r_out) {
  return guarded<true>(ctx, [&] {
    REQUIRE(r_out, "r_out is NULL");
    int st = ctx->read_flags();
    *r_out = ctx->r;
    return st;
  });
}

int lri_snapshots_reserve(lri_ctx* ctx, int64_t count, int dtype) {
  return guarded(ctx, [&] {
    REQUIRE(count >= 1, "count must be >= 1");
    REQUIRE(dtype == LRI_F64 || dtype == LRI_F32 || dtype == LRI_BF16, "unknown dtype");
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (void* p : ctx->owned) HIPCHECK(hipFree(p));
    ctx->owned.clear();
    ctx->snaps.clear();
    ctx->sdtype = dtype;
    // BF16: a column pitch that is a multiple of 4 KB makes some pairs of
    // snapshots stream markedly slower in the passes (which read two at the same
    // offsets at once); 64 elements of padding, never read, remove it.  (The
    // FP32 passes measured 3 % slower with it, the FP64 ones the same.)
    const int64_t es = (int64_t)dtype_size(dtype);
    ctx->sld = ctx->n + ((es == 2 && (ctx->n * es) % 4096 == 0) ? 64 : 0);
    const size_t bytes = dtype_size(dtype) * ctx->sld * ctx->m;
    for (int64_t k = 0; k < count; ++k) {
      void* p = nullptr;
      HIPCHECK(hipMalloc(&p, bytes));
      ctx->owned.push_back(p);
      ctx->snaps.push_back(p);
    }
    ctx->pkN_mode = ctx->pkT_mode = -1;
    return LRI_OK;
  });
}

int lri_snapshot_upload(lri_ctx* ctx, int64_t k, const void* host, int64_t ld) {
  return guarded(ctx, [&] {
    REQUIRE(!ctx->owned.empty(), "no library-owned snapshot slots (call lri_snapshots_reserve)");
    REQUIRE(k >= 0 && k < (int64_t)ctx->owned.size(), "slot out of range");
    REQUIRE(host && ld >= ctx->n, "bad host pointer or leading dimension");
    const size_t es = dtype_size(ctx->sdtype);
    HIPCHECK(hipMemcpy2DAsync(ctx->owned[k], es * ctx->sld, host, es * ld, es * ctx->n, ctx->m,
                              hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return LRI_OK;
  });
}

int lri_snapshots_bind(lri_ctx* ctx, const void* const* dev_ptrs, int64_t count, int dtype, int64_t ld) {
  return guarded(ctx, [&] {
    REQUIRE(dev_ptrs && count >= 1, "need count >= 1 device pointers");
    REQUIRE(dtype == LRI_F64 || dtype == LRI_F32 || dtype == LRI_BF16, "unknown dtype");
    REQUIRE(ld >= ctx->n, "ld must be >= n");
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (void* p : ctx->owned) HIPCHECK(hipFree(p));
    ctx->owned.clear();
    ctx->snaps.assign(dev_ptrs, dev_ptrs + count);
    for (auto p : ctx->snaps) REQUIRE(p != nullptr, "NULL snapshot pointer");
    ctx->sdtype = dtype;
    ctx->sld = ld;
    ctx->pkN_mode = ctx->pkT_mode = -1;
    return LRI_OK;
  });
}

int lri_step_ksl(lri_ctx* ctx, int order, int64_t k_prev, int64_t k_next) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    REQUIRE(order == LRI_PRIMAL || order == LRI_DUAL,
            "lri_step_ksl takes LRI_PRIMAL or LRI_DUAL (Strang: lri_step_ksl_strang)");
    ctx->check_step(k_prev, k_next);
    if (order == LRI_PRIMAL) ctx->ksl_primal(k_prev, k_next);
    else ctx->ksl_dual(k_prev, k_next);
    return LRI_OK;
  });
}

int lri_step_ksl_strang(lri_ctx* ctx, int64_t k_prev, int64_t k_mid, int64_t k_next) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_prev, k_mid);
    ctx->check_step(k_mid, k_next);
    ctx->ksl_primal(k_prev, k_mid);
    ctx->ksl_dual(k_mid, k_next);
    return LRI_OK;
  });
}

int lri_step_bug(lri_ctx* ctx, int64_t k_prev, int64_t k_next) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_prev, k_next);
    ctx->bug(k_prev, k_next);
    return LRI_OK;
  });
}

namespace {
// the host side of the reverse steps: cotangents in, step, cotangents out; the
// KSL step keeps dA V0 in P[:, :r], so its cotangent of U sits in P[:, r:2r]
void data_adjoint_checks(lri_ctx* ctx, double* Ub, int64_t ldub, double* Sb, int64_t ldsb, double* Vb, int64_t ldvb) {
  REQUIRE(Ub && Sb && Vb, "Ub, Sb and Vb are required");
  REQUIRE(ldub >= ctx->n && ldvb >= ctx->m && ldsb >= ctx->r, "leading dimensions");
  if (!ctx->adj) {
    HIPCHECK(hipMalloc((void**)&ctx->adj, sizeof(BugAdjSmall)));
    HIPCHECK(hipMemset(ctx->adj, 0, sizeof(BugAdjSmall)));
  }
}

int adjoint_entry(lri_ctx* ctx, bool ksl, int64_t k_prev, int64_t k_next, double* Ub, int64_t ldub, double* Sb,
                  int64_t ldsb, double* Vb, int64_t ldvb) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_prev, k_next);
    data_adjoint_checks(ctx, Ub, ldub, Sb, ldsb, Vb, ldvb);
    ctx->data_adjoint(ksl, k_prev, k_next, Ub, ldub, Sb, ldsb, Vb, ldvb);
    return ctx->read_flags();
  });
}
}  // namespace

int lri_step_bug_adjoint(lri_ctx* ctx, int64_t k_prev, int64_t k_next, double* Ub, int64_t ldub, double* Sb,
                         int64_t ldsb, double* Vb, int64_t ldvb) {
  return adjoint_entry(ctx, false, k_prev, k_next, Ub, ldub, Sb, ldsb, Vb, ldvb);
}

int lri_step_ksl_adjoint(lri_ctx* ctx, int64_t k_prev, int64_t k_next, double* Ub, int64_t ldub, double* Sb,
                         int64_t ldsb, double* Vb, int64_t ldvb) {
  return adjoint_entry(ctx, true, k_prev, k_next, Ub, ldub, Sb, ldsb, Vb, ldvb);
}

int lri_step_ksl_dual_adjoint(lri_ctx* ctx, int64_t k_prev, int64_t k_next, double* Ub, int64_t ldub, double* Sb,
                              int64_t ldsb, double* Vb, int64_t ldvb) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_prev, k_next);
    data_adjoint_checks(ctx, Ub, ldub, Sb, ldsb, Vb, ldvb);
    ctx->ksl_dual_adjoint(k_prev, k_next, Ub, ldub, Sb, ldsb, Vb, ldvb);
    return ctx->read_flags();
  });
}

int lri_step_ksl_strang_adjoint(lri_ctx* ctx, int64_t k_prev, int64_t k_mid, int64_t k_next, double* Ub, int64_t ldub,
                                double* Sb, int64_t ldsb, double* Vb, int64_t ldvb) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_prev, k_mid);
    ctx->check_step(k_mid, k_next);
    data_adjoint_checks(ctx, Ub, ldub, Sb, ldsb, Vb, ldvb);
    // the Strang step is primal(A[k_mid] - A[k_prev]) then dual(A[k_next] - A[k_mid])
    // (projector_splitting.jl:205-211): the state after the first half is
    // recomputed from a copy of the step's start, the dual half reversed there,
    // then the primal half reversed at the start
    ctx->save_state();
    ctx->ksl_primal(k_prev, k_mid);
    const int st = ctx->read_flags();
    if (st != LRI_OK) {
      ctx->restore_state();
      return st;
    }
    try {
      ctx->ksl_dual_adjoint(k_mid, k_next, Ub, ldub, Sb, ldsb, Vb, ldvb);
    } catch (...) {
      ctx->restore_state();
      throw;
    }
    ctx->restore_state();
    ctx->data_adjoint(true, k_prev, k_mid, Ub, ldub, Sb, ldsb, Vb, ldvb);
    return ctx->read_flags();
  });
}

int lri_step_bug_ra_adjoint(lri_ctx* ctx, int64_t k_prev, int64_t k_next, double tol, int64_t rmax, int64_t r_new,
                            double* Ub, int64_t ldub, double* Sb, int64_t ldsb, double* Vb, int64_t ldvb) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_prev, k_next);
    REQUIRE(tol >= 0.0, "tol must be >= 0");
    REQUIRE(rmax >= 1, "rmax must be >= 1");
    REQUIRE(Ub && Sb && Vb, "Ub, Sb and Vb are required");
    const int64_t r = ctx->r, n = ctx->n, m = ctx->m;
    REQUIRE(r_new >= 1 && r_new <= std::min<int64_t>(2 * r, ctx->rcap) && r_new <= rmax, "r_new out of range");
    const int64_t rw = std::max(r, r_new);
    REQUIRE(ldub >= n && ldvb >= m && ldsb >= rw, "leading dimensions");
    if (!ctx->radj) {
      HIPCHECK(hipMalloc((void**)&ctx->radj, sizeof(RaAdjSmall)));
      HIPCHECK(hipMemset(ctx->radj, 0, sizeof(RaAdjSmall)));
    }
    const size_t d = sizeof(double);
    HIPCHECK(hipMemcpy2DAsync(ctx->P, d * n, Ub, d * ldub, d * n, r_new, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(ctx->Q, d * m, Vb, d * ldvb, d * m, r_new, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(ctx->radj->Sb1, d * r_new, Sb, d * ldsb, d * r_new, r_new, hipMemcpyDefault,
                              ctx->stream));
    ctx->bug_ra_adjoint(k_prev, k_next, tol, rmax, (int)r_new);
    HIPCHECK(hipMemcpy2DAsync(Ub, d * ldub, ctx->P, d * n, d * n, r, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(Vb, d * ldvb, ctx->Q, d * m, d * m, r, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(Sb, d * ldsb, ctx->radj->Sb0, d * r, d * r, r, hipMemcpyDefault, ctx->stream));
    return ctx->read_flags();
  });
}

int lri_step_greedy(lri_ctx* ctx, int64_t k_next) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_next, k_next);
    ctx->greedy(k_next);
    return LRI_OK;
  });
}

int lri_step_greedy_twofactor(lri_ctx* ctx, int64_t k_next) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_next, k_next);
    ctx->greedy_twofactor(k_next);
    return LRI_OK;
  });
}

int lri_step_bug_ra(lri_ctx* ctx, int64_t k_prev, int64_t k_next, double tol, int64_t rmax, int64_t* r_new) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_prev, k_next);
    REQUIRE(tol >= 0.0, "tol must be >= 0");
    REQUIRE(rmax >= 1, "rmax must be >= 1");
    int64_t r1 = ctx->bug_ra(k_prev, k_next, tol, rmax);
    if (r_new) *r_new = r1;
    return ctx->read_flags();
  });
}

namespace {

// A matrix-ODE step does not pack the data-path panels, and it leaves the ones
// packed before stale.
struct DeScope {
  lri_ctx* c;
  explicit DeScope(lri_ctx* c_) : c(c_) { c->pack_enabled = false; }
  ~DeScope() {
    c->pack_enabled = true;
    c->pkN_mode = c->pkT_mode = -1;
  }
};

}  // namespace

int lri_set_operator_lyapunov(lri_ctx* ctx, const double* a_diag, const double* a_off, int a_periodic,
                              const double* b_diag, const double* b_off, int b_periodic, int64_t q,
                              const double* G, int64_t ldg, const double* H, int64_t ldh) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    REQUIRE(a_diag && b_diag, "a_diag and b_diag are required");
    REQUIRE((a_off || ctx->n == 1) && (b_off || ctx->m == 1), "a_off and b_off are required");
    REQUIRE(q >= 0 && q <= LRI_MAX_FORCE, "forcing rank q must be in [0, LRI_MAX_FORCE]");
    REQUIRE(q == 0 || (G && H && ldg >= ctx->n && ldh >= ctx->m), "G, H or their leading dimensions");
    REQUIRE(!a_periodic || ctx->n >= 3, "a periodic operator needs n >= 3");
    REQUIRE(!b_periodic || ctx->m >= 3, "a periodic operator needs m >= 3");
    REQUIRE(ctx->n <= LRI_DE_MAX_ROWS && ctx->m <= LRI_DE_MAX_ROWS, "matrix ODE needs n, m <= LRI_DE_MAX_ROWS");
    const size_t d = sizeof(double);
    const int64_t n = ctx->n, m = ctx->m;
    ctx->drop_graphs();
    auto alloc = [&](double** p, size_t count) {
      if (!*p) HIPCHECK(hipMalloc((void**)p, d * count));
      HIPCHECK(hipMemset(*p, 0, d * count));
    };
    alloc(&ctx->op_ad, n);
    alloc(&ctx->op_ae, n);
    alloc(&ctx->op_bd, m);
    alloc(&ctx->op_be, m);
    alloc(&ctx->op_G, n * LRI_MAX_FORCE);
    alloc(&ctx->op_H, m * LRI_MAX_FORCE);
    if (!ctx->de) {
      HIPCHECK(hipMalloc((void**)&ctx->de, sizeof(DeSmall)));
      HIPCHECK(hipMemset(ctx->de, 0, sizeof(DeSmall)));
    }
    HIPCHECK(hipMemcpy(ctx->op_ad, a_diag, d * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(ctx->op_bd, b_diag, d * m, hipMemcpyHostToDevice));
    // a_off: n entries when periodic, n - 1 otherwise (the last stays zero)
    const int64_t na = a_periodic ? n : n - 1, nb = b_periodic ? m : m - 1;
    if (na > 0) HIPCHECK(hipMemcpy(ctx->op_ae, a_off, d * na, hipMemcpyHostToDevice));
    if (nb > 0) HIPCHECK(hipMemcpy(ctx->op_be, b_off, d * nb, hipMemcpyHostToDevice));
    if (q > 0) {
      HIPCHECK(hipMemcpy2D(ctx->op_G, d * n, G, d * ldg, d * n, q, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy2D(ctx->op_H, d * m, H, d * ldh, d * m, q, hipMemcpyHostToDevice));
    }
    ctx->op_aper = a_periodic != 0;
    ctx->op_bper = b_periodic != 0;
    ctx->op_q = (int)q;
    ctx->setup_spectral(a_diag, a_off, a_periodic != 0, b_diag, b_off, b_periodic != 0);
    ctx->op_quad = false;
    ctx->op_set = true;
    return LRI_OK;
  });
}

int lri_set_operator_quadratic(lri_ctx* ctx, const double* dl, const double* dd, const double* du, int periodic) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    REQUIRE(ctx->op_set, "set the linear part first (lri_set_operator_lyapunov)");
    REQUIRE(dl && dd && du, "dl, dd and du are required (n entries each)");
    REQUIRE(!periodic || ctx->n >= 3, "a periodic operator needs n >= 3");
    const size_t d = sizeof(double);
    const int64_t n = ctx->n;
    ctx->drop_graphs();
    auto alloc = [&](double** p, size_t count) {
      if (!*p) HIPCHECK(hipMalloc((void**)p, d * count));
      HIPCHECK(hipMemset(*p, 0, d * count));
    };
    const size_t k3 = (size_t)kQuadMaxRank * kQuadMaxRank * kQuadMaxRank;
    const int64_t rows = std::max(ctx->n, ctx->m);
    alloc(&ctx->op_dl, n);
    alloc(&ctx->op_dd, n);
    alloc(&ctx->op_du, n);
    alloc(&ctx->quad_part, (size_t)((rows + 63) / 64) * k3);
    alloc(&ctx->quad_T, k3);
    alloc(&ctx->quad_W, k3);
    alloc(&ctx->quad_work, 2 * k3);
    HIPCHECK(hipMemcpy(ctx->op_dl, dl, d * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(ctx->op_dd, dd, d * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(ctx->op_du, du, d * n, hipMemcpyHostToDevice));
    ctx->op_dper = periodic != 0;
    ctx->op_quad = true;
    return LRI_OK;
  });
}

int lri_step_hybrid_greedy(lri_ctx* ctx, int64_t k_next, double dt, int64_t substeps) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_next, k_next);
    ctx->check_de(dt, substeps);
    ctx->hybrid_greedy(k_next, dt, substeps);
    return LRI_OK;
  });
}

int lri_step_de_ksl(lri_ctx* ctx, int order, double dt, int64_t substeps) {
  return guarded(ctx, [&] {
    REQUIRE(order == LRI_PRIMAL || order == LRI_DUAL || order == LRI_STRANG, "unknown order");
    ctx->check_de(dt, substeps);
    DeScope scope(ctx);
    if (order == LRI_PRIMAL) {
      ctx->de_ksl_primal(dt, substeps);
    } else if (order == LRI_DUAL) {
      ctx->de_ksl_dual(dt, substeps);
    } else {
      ctx->de_ksl_primal(0.5 * dt, substeps);
      ctx->de_ksl_dual(0.5 * dt, substeps);
    }
    return LRI_OK;
  });
}

int lri_step_de_bug(lri_ctx* ctx, double dt, int64_t substeps) {
  return guarded<true>(ctx, [&] {
    ctx->check_de(dt, substeps);
    DeScope scope(ctx);
    ctx->de_bug(dt, substeps);
    return LRI_OK;
  });
}

static int de_adjoint_entry(lri_ctx* ctx, bool ksl, double dt, int64_t substeps, double* Ub, int64_t ldub,
                            double* Sb, int64_t ldsb, double* Vb, int64_t ldvb, double* scale_bar) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_de(dt, substeps);
    REQUIRE(ctx->op_quad,
            "the matrix-ODE adjoint needs the quadratic operator (lri_set_operator_quadratic; D = 0 for a "
            "linear right-hand side)");
    REQUIRE(Ub && Sb && Vb, "Ub, Sb and Vb are required");
    const int64_t r = ctx->r, n = ctx->n, m = ctx->m;
    REQUIRE(ldub >= n && ldvb >= m && ldsb >= r, "leading dimensions");
    DeScope scope(ctx);
    if (!ctx->adj) {
      HIPCHECK(hipMalloc((void**)&ctx->adj, sizeof(BugAdjSmall)));
      HIPCHECK(hipMemset(ctx->adj, 0, sizeof(BugAdjSmall)));
    }
    const size_t d = sizeof(double);
    HIPCHECK(hipMemcpy2DAsync(ctx->P, d * n, Ub, d * ldub, d * n, r, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(ctx->Q, d * m, Vb, d * ldvb, d * m, r, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(ctx->adj->Sb1, d * r, Sb, d * ldsb, d * r, r, hipMemcpyDefault, ctx->stream));
    if (ksl)
      ctx->quad_ksl_adjoint(dt, substeps);
    else
      ctx->quad_bug_adjoint(dt, substeps);
    double sc[2] = {0.0, 0.0};
    HIPCHECK(hipMemcpy2DAsync(Ub, d * ldub, ctx->P, d * n, d * n, r, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(Vb, d * ldvb, ctx->Q, d * m, d * m, r, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(Sb, d * ldsb, ctx->adj->Sb0, d * r, d * r, r, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpyAsync(sc, ctx->qadj->scal, sizeof(sc), hipMemcpyDefault, ctx->stream));
    const int f = ctx->read_flags();
    if (scale_bar) {
      scale_bar[0] += sc[0];
      scale_bar[1] += sc[1];
    }
    return f;
  });
}

int lri_step_de_bug_adjoint(lri_ctx* ctx, double dt, int64_t substeps, double* Ub, int64_t ldub, double* Sb,
                            int64_t ldsb, double* Vb, int64_t ldvb, double* scale_bar) {
  return de_adjoint_entry(ctx, false, dt, substeps, Ub, ldub, Sb, ldsb, Vb, ldvb, scale_bar);
}

int lri_step_de_ksl_adjoint(lri_ctx* ctx, double dt, int64_t substeps, double* Ub, int64_t ldub, double* Sb,
                            int64_t ldsb, double* Vb, int64_t ldvb, double* scale_bar) {
  return de_adjoint_entry(ctx, true, dt, substeps, Ub, ldub, Sb, ldsb, Vb, ldvb, scale_bar);
}

int lri_step_de_bug_ra_adjoint(lri_ctx* ctx, double dt, int64_t substeps, double tol, int64_t rmax, int64_t r_new,
                               double* Ub, int64_t ldub, double* Sb, int64_t ldsb, double* Vb, int64_t ldvb,
                               double* scale_bar) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_de(dt, substeps);
    REQUIRE(ctx->op_quad,
            "the matrix-ODE adjoint needs the quadratic operator (lri_set_operator_quadratic; D = 0 for a "
            "linear right-hand side)");
    REQUIRE(tol >= 0.0, "tol must be >= 0");
    REQUIRE(rmax >= 1, "rmax must be >= 1");
    REQUIRE(Ub && Sb && Vb, "Ub, Sb and Vb are required");
    const int64_t r = ctx->r, n = ctx->n, m = ctx->m;
    REQUIRE(2 * r <= kQuadMaxRank, "the rank-adaptive step with the quadratic operator needs 2 r <= LRI_QUAD_MAX_RANK");
    REQUIRE(r_new >= 1 && r_new <= std::min<int64_t>(2 * r, ctx->rcap) && r_new <= rmax, "r_new out of range");
    const int64_t rw = std::max(r, r_new);
    REQUIRE(ldub >= n && ldvb >= m && ldsb >= rw, "leading dimensions");
    DeScope scope(ctx);
    if (!ctx->radj) {
      HIPCHECK(hipMalloc((void**)&ctx->radj, sizeof(RaAdjSmall)));
      HIPCHECK(hipMemset(ctx->radj, 0, sizeof(RaAdjSmall)));
    }
    if (!ctx->adj) {
      HIPCHECK(hipMalloc((void**)&ctx->adj, sizeof(BugAdjSmall)));
      HIPCHECK(hipMemset(ctx->adj, 0, sizeof(BugAdjSmall)));
    }
    const size_t d = sizeof(double);
    HIPCHECK(hipMemcpy2DAsync(ctx->P, d * n, Ub, d * ldub, d * n, r_new, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(ctx->Q, d * m, Vb, d * ldvb, d * m, r_new, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(ctx->radj->Sb1, d * r_new, Sb, d * ldsb, d * r_new, r_new, hipMemcpyDefault,
                              ctx->stream));
    ctx->quad_bug_ra_adjoint(dt, substeps, tol, rmax, (int)r_new);
    double sc[2] = {0.0, 0.0};
    HIPCHECK(hipMemcpy2DAsync(Ub, d * ldub, ctx->P, d * n, d * n, r, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(Vb, d * ldvb, ctx->Q, d * m, d * m, r, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpy2DAsync(Sb, d * ldsb, ctx->adj->Sb0, d * r, d * r, r, hipMemcpyDefault, ctx->stream));
    HIPCHECK(hipMemcpyAsync(sc, ctx->qadj->scal, sizeof(sc), hipMemcpyDefault, ctx->stream));
    const int f = ctx->read_flags();
    if (scale_bar) {
      scale_bar[0] += sc[0];
      scale_bar[1] += sc[1];
    }
    return f;
  });
}

int lri_step_de_bug_ra(lri_ctx* ctx, double dt, int64_t substeps, double tol, int64_t rmax, int64_t* r_new) {
  return guarded(ctx, [&] {
    ctx->check_de(dt, substeps);
    REQUIRE(tol >= 0.0, "tol must be >= 0");
    REQUIRE(rmax >= 1, "rmax must be >= 1");
    REQUIRE(!ctx->op_quad || 2 * ctx->r <= kQuadMaxRank,
            "the rank-adaptive step with the quadratic operator needs 2 r <= LRI_QUAD_MAX_RANK");
    int64_t r1 = 0;
    {
      DeScope scope(ctx);
      r1 = ctx->de_bug_ra(dt, substeps, tol, rmax);
    }
    if (r_new) *r_new = r1;
    return ctx->read_flags();
  });
}


int lri_batch_create(lri_batch** out, int device, int64_t count, const int64_t* n, const int64_t* m,
                     const int64_t* rcap) {
  g_create_error.clear();
  if (!out || count < 1 || !n || !m || !rcap) {
    g_create_error = "out, n, m and rcap are required and count >= 1";
    return LRI_EINVAL;
  }
  *out = nullptr;
  for (int64_t i = 0; i < count; ++i)  // every member's shape before any device work
    if (!shape_ok(n[i], m[i], rcap[i])) {
      g_create_error = "problem " + std::to_string(i) + ": need n, m >= 1 and 1 <= rcap <= min(n, m, LRI_MAX_RANK)";
      return LRI_EINVAL;
    }
  lri_batch* b = nullptr;
  try {
    b = new lri_batch();
  } catch (const std::bad_alloc&) {
    g_create_error = "host allocation failed";
    return LRI_EOOM;
  }
  b->device = device;
  for (int64_t i = 0; i < count; ++i) {
    lri_ctx* c = nullptr;
    const int st = lri_create(&c, device, n[i], m[i], rcap[i]);
    if (st != LRI_OK) {
      g_create_error = "problem " + std::to_string(i) + ": " + g_create_error;
      batch_free(b);
      return st;
    }
    b->ctx.push_back(c);
  }
  try {
    batch_set_merged(b, true);
  } catch (const Failure& e) {
    g_create_error = e.msg;
    batch_free(b);
    return e.status;
  } catch (const std::bad_alloc&) {
    g_create_error = "host allocation failed";
    batch_free(b);
    return LRI_EOOM;
  }
  *out = b;
  return LRI_OK;
}

int lri_batch_destroy(lri_batch* b) {
  if (!b) return LRI_OK;
  batch_free(b);
  return LRI_OK;
}

const char* lri_batch_last_error(const lri_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int lri_batch_context(lri_batch* b, int64_t i, lri_ctx** ctx) {
  if (!b || !ctx || i < 0 || i >= (int64_t)b->ctx.size()) return LRI_EINVAL;
  *ctx = b->ctx[i];
  return LRI_OK;
}


int lri_batch_step_de_bug(lri_batch* b, double dt, const int64_t* substeps) {
  if (!substeps) return LRI_EINVAL;
  return batch_each(b, [&](int64_t i, lri_ctx* c) { return lri_step_de_bug(c, dt, substeps[i]); });
}

namespace {

// Every member can run as merged launches: its step records completely (the
// resident eigenbasis, no quadratic term, no per-launch timing).
bool batch_mergeable(const lri_batch* b) {
  if (!b->merged) return false;
  for (const lri_ctx* c : b->ctx)
    if (!c->resident || c->op_quad || c->timing) return false;
  return true;
}

// lri_batch_step_de_bug_ra as merged launches: each member's step up to its SVD
// (phase 1, recorded per member and reused while the member's rank, the step
// parameters and its epoch stay), one gather of every member's new rank and
// condition flags and one wait, then each member's rotation to its new rank
// (phase 2).  A member whose phase 1 raised a condition, or whose new rank is
// above its capacity, keeps its state and reports the status.  Phase 2 is left
// queued on the shared stream.
int merged_step_de_bug_ra(lri_batch* b, double dt, const int64_t* substeps, double tol, int64_t rmax,
                          int64_t* r_new) {
  const int count = (int)b->ctx.size();
  b->status.assign(count, LRI_OK);
  int first = LRI_OK;
  auto fail = [&](int i, int st) {
    b->status[i] = st;
    if (first == LRI_OK) {
      first = st;
      b->err = "problem " + std::to_string(i) + ": " + lri_last_error(b->ctx[i]);
    }
  };
  auto device_failure = [&](const Failure& e) {
    b->err = e.msg;
    for (int i = 0; i < count; ++i)
      if (b->status[i] == LRI_OK) b->status[i] = e.status;
    return e.status;
  };
  try {
    HIPCHECK(hipSetDevice(b->device));
  } catch (const Failure& e) {
    return device_failure(e);
  }
  for (int i = 0; i < count; ++i) {
    lri_ctx* c = b->ctx[i];
    b->ph1.part[i] = 0;
    const int st = guarded(c, [&] {
      c->check_de(dt, substeps[i]);
      REQUIRE(tol >= 0.0, "tol must be >= 0");
      REQUIRE(rmax >= 1, "rmax must be >= 1");
      const int64_t rm = std::min<int64_t>(rmax, c->rcap);
      const MergedPhase::Key key{(double)c->r, (double)substeps[i], dt, tol, (double)rm, (double)c->epoch, 0.0, 0.0};
      b->ph1.record(i, key, [&] {
        DeScope scope(c);
        c->defer_rank_copy = true;
        try {
          c->de_bug_ra(dt, substeps[i], tol, rm, false);
        } catch (...) {
          c->defer_rank_copy = false;
          throw;
        }
        c->defer_rank_copy = false;
      });
      c->ra_k2 = 2 * (int)c->r;
      return LRI_OK;
    });
    if (st != LRI_OK) fail(i, st);
    else b->ph1.part[i] = 1;
  }
  try {
    b->ph1.upload(b->stream);
    b->ph1.run(b->stream, b->graphs);
    hipLaunchKernelGGL(k_gather_status, dim3((count + 255) / 256), dim3(256), 0, b->stream, b->words, b->word_out,
                       count);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(b->word_host, b->word_out, sizeof(int) * 2 * count, hipMemcpyDeviceToHost, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
  } catch (const Failure& e) {
    return device_failure(e);
  }
  for (int i = 0; i < count; ++i) {
    b->ph2.part[i] = 0;
    if (!b->ph1.part[i]) continue;
    lri_ctx* c = b->ctx[i];
    const int r1 = b->word_host[i], f = b->word_host[count + i];
    const int st = guarded(c, [&] {
      c->pkN_mode = c->pkT_mode = -1;
      if (f & FLAG_NONFINITE) {
        c->err = "non-finite values reached a Gram matrix; the state is left at the step's start";
        return LRI_ENONFINITE;
      }
      if (f & FLAG_BREAKDOWN) {
        c->err = "numerically singular Gram matrix; the state is left at the step's start";
        return LRI_EBREAKDOWN;
      }
      c->ra_check_rank(r1);
      const int k2 = 2 * (int)c->r;
      const MergedPhase::Key key{(double)k2, (double)r1, (double)c->epoch, 0.0, 0.0, 0.0, 0.0, 0.0};
      b->ph2.record(i, key, [&] {
        DeScope scope(c);
        c->ra_apply(k2, r1);
      });
      c->r = r1;
      if (r_new) r_new[i] = r1;
      return LRI_OK;
    });
    if (st != LRI_OK) fail(i, st);
    else b->ph2.part[i] = 1;
  }
  try {
    b->ph2.upload(b->stream);
    b->ph2.run(b->stream, b->graphs);
  } catch (const Failure& e) {
    return device_failure(e);
  }
  ++b->merged_steps;
  return first;
}

}  // namespace

int lri_batch_step_de_bug_ra(lri_batch* b, double dt, const int64_t* substeps, double tol, int64_t rmax,
                             int64_t* r_new) {
  if (!b || !substeps) return LRI_EINVAL;
  if (batch_mergeable(b)) return merged_step_de_bug_ra(b, dt, substeps, tol, rmax, r_new);
  // queue every member's step up to its SVD ...
  int st = batch_each(b, [&](int64_t i, lri_ctx* c) {
    return guarded(c, [&] {
      c->check_de(dt, substeps[i]);
      REQUIRE(tol >= 0.0, "tol must be >= 0");
      REQUIRE(rmax >= 1, "rmax must be >= 1");
      REQUIRE(!c->op_quad || 2 * c->r <= kQuadMaxRank,
              "the rank-adaptive step with the quadratic operator needs 2 r <= LRI_QUAD_MAX_RANK");
      DeScope scope(c);
      // a member's capacity is its own rmax (the reference gives each problem its own
      // algorithm, rank_adaptive_unconventional.jl:18-23)
      const int64_t rm = std::min<int64_t>(rmax, c->rcap);
      if (b->graphs && !c->timing)
        c->de_bug_ra_queue_graph(dt, substeps[i], tol, rm);
      else
        c->de_bug_ra(dt, substeps[i], tol, rm, false);
      return LRI_OK;
    });
  });
  // ... then wait for the ranks in order and queue each member's rotation
  // (members whose first half failed are skipped and keep their state) ...
  const int st2 = batch_each(
      b,
      [&](int64_t i, lri_ctx* c) {
        return guarded(c, [&] {
          DeScope scope(c);
          const int64_t r1 = c->ra_rotate();
          if (r_new) r_new[i] = r1;
          return LRI_OK;
        });
      },
      false);
  // ... and collect (and clear) the device-side conditions of every member once
  // all are queued, including those of members that failed above
  std::vector<int> before = b->status;
  std::fill(b->status.begin(), b->status.end(), LRI_OK);
  const int st3 = batch_each(b, [&](int64_t, lri_ctx* c) { return guarded<true>(c, [&] { return c->read_flags(); }); },
                             false);
  for (size_t i = 0; i < before.size(); ++i)
    if (before[i] != LRI_OK) b->status[i] = before[i];
  if (st != LRI_OK) return st;
  if (st2 != LRI_OK) return st2;
  return st3;
}

int lri_batch_member_status(const lri_batch* b, int64_t i, int* status) {
  if (!b || !status || i < 0 || i >= (int64_t)b->ctx.size()) return LRI_EINVAL;
  *status = i < (int64_t)b->status.size() ? b->status[i] : LRI_OK;
  return LRI_OK;
}

int lri_batch_use_graphs(lri_batch* b, int on) {
  if (!b) return LRI_EINVAL;
  b->graphs = on != 0;
  return LRI_OK;
}

int lri_batch_use_merged_launches(lri_batch* b, int on) {
  if (!b) return LRI_EINVAL;
  try {
    batch_set_merged(b, on != 0);
  } catch (const Failure& e) {
    b->err = e.msg;
    return e.status;
  }
  return LRI_OK;
}

int lri_batch_launch_stats(const lri_batch* b, int64_t* merged_steps, int64_t* launches) {
  if (!b) return LRI_EINVAL;
  if (merged_steps) *merged_steps = b->merged_steps;
  if (launches) *launches = b->merged_steps ? (int64_t)(b->ph1.launches.size() + 1 + b->ph2.launches.size()) : 0;
  return LRI_OK;
}

int lri_batch_synchronize(lri_batch* b) {
  return batch_each(b, [&](int64_t, lri_ctx* c) { return lri_synchronize(c); });
}

int lri_set_resident_eigenbasis(lri_ctx* ctx, int on) {
  return guarded(ctx, [&] {
    if (on) ctx->enter_resident();
    else ctx->leave_resident();
    return LRI_OK;
  });
}

int lri_trajectory_reserve(lri_ctx* ctx, int64_t slots) {
  return guarded<true>(ctx, [&] {
    REQUIRE(slots >= 0, "slots must be >= 0");
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->traj) {
      HIPCHECK(hipFree(ctx->traj));
      ctx->traj = nullptr;
    }
    ctx->traj_slots = 0;
    ctx->traj_rank.assign((size_t)slots, 0);
    ctx->traj_resident.assign((size_t)slots, 0);
    if (slots > 0) {
      HIPCHECK(hipMalloc((void**)&ctx->traj, sizeof(double) * (size_t)slots * ctx->traj_stride()));
      ctx->traj_slots = slots;
    }
    return LRI_OK;
  });
}

int lri_trajectory_save(lri_ctx* ctx, int64_t slot) {
  return guarded<true>(ctx, [&] {
    REQUIRE(slot >= 0 && slot < ctx->traj_slots, "trajectory slot out of range (lri_trajectory_reserve)");
    REQUIRE(ctx->r >= 1, "state not set (call lri_set_state)");
    const size_t d = sizeof(double);
    const int64_t n = ctx->n, m = ctx->m, rc = ctx->rcap, r = ctx->r;
    double* base = ctx->traj + slot * ctx->traj_stride();
    HIPCHECK(hipMemcpyAsync(base, ctx->U, d * n * r, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(hipMemcpyAsync(base + n * rc, ctx->S, d * rc * rc, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(hipMemcpyAsync(base + n * rc + rc * rc, ctx->V, d * m * r, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->traj_rank[(size_t)slot] = r;
    ctx->traj_resident[(size_t)slot] = ctx->resident ? 1 : 0;
    return LRI_OK;
  });
}

int lri_trajectory_get(lri_ctx* ctx, int64_t slot, int64_t* r_out, double* U, int64_t ldu, double* S, int64_t lds,
                       double* V, int64_t ldv) {
  return guarded<true>(ctx, [&] {
    REQUIRE(slot >= 0 && slot < ctx->traj_slots, "trajectory slot out of range (lri_trajectory_reserve)");
    const int64_t r = ctx->traj_rank[(size_t)slot];
    REQUIRE(r >= 1, "trajectory slot was never saved");
    const int64_t n = ctx->n, m = ctx->m, rc = ctx->rcap;
    REQUIRE((!U || ldu >= n) && (!S || lds >= r) && (!V || ldv >= m), "leading dimensions too small");
    const size_t d = sizeof(double);
    const double* base = ctx->traj + slot * ctx->traj_stride();
    const double* Us = base;
    const double* Vs = base + n * rc + rc * rc;
    if (ctx->traj_resident[(size_t)slot] && (U || V)) {  // saved in the eigenbasis
      REQUIRE(ctx->resident, "the slot was saved in the resident eigenbasis, which has been left since");
      ctx->resident_out(Us, Vs, (int)r);
      Us = ctx->K2;
      Vs = ctx->L2;
    }
    int st = ctx->read_flags();
    if (U) HIPCHECK(hipMemcpy2D(U, d * ldu, Us, d * n, d * n, r, hipMemcpyDefault));
    if (S) HIPCHECK(hipMemcpy2D(S, d * lds, base + n * rc, d * rc, d * r, r, hipMemcpyDefault));
    if (V) HIPCHECK(hipMemcpy2D(V, d * ldv, Vs, d * m, d * m, r, hipMemcpyDefault));
    if (r_out) *r_out = r;
    return st;
  });
}

int lri_trajectory_get_range(lri_ctx* ctx, int64_t slot0, int64_t count, int64_t* ranks, double* host) {
  return guarded<true>(ctx, [&] {
    REQUIRE(count >= 0 && slot0 >= 0 && slot0 + count <= ctx->traj_slots,
            "trajectory slots out of range (lri_trajectory_reserve)");
    REQUIRE(count == 0 || (ranks && host), "ranks and host are required");
    const int64_t stride = ctx->traj_stride();
    for (int64_t i = 0; i < count; ++i) {
      const size_t k = (size_t)(slot0 + i);
      REQUIRE(ctx->traj_rank[k] >= 1, "a trajectory slot of the range was never saved");
      REQUIRE(!ctx->traj_resident[k], "a slot of the range was saved in the resident eigenbasis: use lri_trajectory_get");
      ranks[i] = ctx->traj_rank[k];
    }
    const int st = ctx->read_flags();
    if (count > 0)
      HIPCHECK(hipMemcpy(host, ctx->traj + slot0 * stride, sizeof(double) * (size_t)(count * stride),
                         hipMemcpyDeviceToHost));
    return st;
  });
}

int lri_synchronize(lri_ctx* ctx) {
  return guarded<true>(ctx, [&] { return ctx->read_flags(); });
}

int lri_get_stream(lri_ctx* ctx, void** stream) {
  return guarded<true>(ctx, [&] {
    REQUIRE(stream, "stream is NULL");
    *stream = (void*)ctx->stream;
    return LRI_OK;
  });
}

int lri_kernel_timing(lri_ctx* ctx, int mode) {
  return guarded<true>(ctx, [&] {
    REQUIRE(mode >= 0 && mode <= 2, "timing mode must be 0, 1 or 2");
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->drain_timing();
    ctx->timing = mode != 0;
    ctx->timing_passes_only = mode == 2;
    for (int f = 0; f < LRI_KERNEL_COUNT; ++f) {
      ctx->tot_ms[f] = 0.0;
      ctx->launches[f] = 0;
      ctx->samples[f].clear();
      ctx->seen[f] = 0;
    }
    return LRI_OK;
  });
}

int lri_autotune_passes(lri_ctx* ctx, int64_t kp, int64_t kn, int64_t reps, double* report) {
  return guarded<true>(ctx, [&] {
    ctx->leave_resident();
    REQUIRE(report, "report is required (8 doubles)");
    REQUIRE(ctx->r >= 1, "state not set (call lri_set_state)");
    REQUIRE(ctx->sdtype >= 0, "no snapshots (call lri_snapshots_reserve)");
    const int64_t ns = (int64_t)ctx->snaps.size();
    REQUIRE(kn >= 0 && kn < ns && ctx->snaps[kn], "snapshot kn is not set");
    REQUIRE(kp < ns && (kp < 0 || ctx->snaps[kp]), "snapshot kp is not set");
    REQUIRE(reps >= 1 && reps <= 1000, "reps must be in [1, 1000]");
    ctx->autotune_passes(kp, kn, (int)reps, report);
    return LRI_OK;
  });
}

int lri_kernel_timing_sample(lri_ctx* ctx, int every) {
  return guarded<true>(ctx, [&] {
    REQUIRE(every >= 1, "every must be >= 1");
    ctx->timing_every = every;
    return LRI_OK;
  });
}

int lri_kernel_timing_samples(lri_ctx* ctx, int family, float* ms, int64_t cap, int64_t* count) {
  return guarded<true>(ctx, [&] {
    REQUIRE(family >= 0 && family < LRI_KERNEL_COUNT, "unknown kernel family");
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->drain_timing();
    const auto& v = ctx->samples[family];
    if (count) *count = (int64_t)v.size();
    if (ms)
      for (int64_t i = 0; i < cap && i < (int64_t)v.size(); ++i) ms[i] = v[i];
    return LRI_OK;
  });
}

int lri_kernel_timing_read(lri_ctx* ctx, int family, double* total_ms, int64_t* launches) {
  return guarded<true>(ctx, [&] {
    REQUIRE(family >= 0 && family < LRI_KERNEL_COUNT, "unknown kernel family");
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->drain_timing();
    if (total_ms) *total_ms = ctx->tot_ms[family];
    if (launches) *launches = ctx->launches[family];
    return LRI_OK;
  });
}

int lri_debug_product(lri_ctx* ctx, int which, int64_t k_prev, int64_t k_next, double* out, int64_t ldo) {
  return guarded(ctx, [&] {
    ctx->leave_resident();
    ctx->check_step(k_prev, k_next);
    REQUIRE(which == 0 || which == 1, "which must be 0 (dA V) or 1 (dA' U)");
    REQUIRE(out != nullptr, "out is NULL");
    const int rr = (int)ctx->r;
    const int64_t rows = which == 0 ? ctx->n : ctx->m;
    REQUIRE(ldo >= rows, "ldo too small");
    ctx->ensure_packed(which == 0);
    void* part = which == 0 ? ctx->partN : ctx->partT;
    const int ns = which == 0 ? ctx->passN(k_prev, k_next, rr, part) : ctx->passT(k_prev, k_next, rr, part);
    double* dst = which == 0 ? ctx->P : ctx->Q;
    ctx->with_parts(part, [&](auto A) {
      A.rows = rows;
      A.nsplit = ns;
      A.rb = lri_ctx::rb_for(ctx->pass_mode(), rr);
      A.r_out = rr;
      A.out = dst;
      A.ldo = rows;
      ctx->rows(A);
    });
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy2D(out, sizeof(double) * ldo, dst, sizeof(double) * rows, sizeof(double) * rows, rr,
                         hipMemcpyDeviceToHost));
    return ctx->read_flags();
  });
}

int lri_selftest_mfma(int which, const void* a, const void* b, void* out) {
  if (which < 0 || which > 2 || !a || !b || !out) return LRI_EINVAL;
  hipError_t e = selftest_mfma(which, a, b, out);
  if (e != hipSuccess) {
    g_create_error = std::string("selftest: ") + hipGetErrorString(e);
    return LRI_EHIP;
  }
  return LRI_OK;
}

int lri_stream_probe(int pattern, double* tbps) {
  if (pattern < 0 || pattern > 2 || !tbps) return LRI_EINVAL;
  hipError_t e = stream_probe(pattern, tbps);
  if (e != hipSuccess) {
    g_create_error = std::string("stream probe: ") + hipGetErrorString(e);
    return LRI_EHIP;
  }
  return LRI_OK;
}

}  // extern "C"
