// Kernel argument blocks and launchers shared by the engine and the kernels.
#pragma once

#include <cstring>
#include <type_traits>
#include <vector>

#include "lri_common.h"

namespace lri {

// ----------------------------------------------------------------------------
// Launch recording, for running the steps of many contexts as merged launches.
// While a Recorder is installed on the calling thread, each launcher below
// appends its launch (kernel, variant, workgroups, threads, LDS and argument
// block) to it instead of launching.  A batch records every member's step with
// the member's own buffers, then runs position p of all the sequences as one
// launch per (kernel, variant, threads) of the kernel's batched form: each
// workgroup finds its member in a table of workgroup offsets and runs the same
// body with the member's arguments and its workgroup index within the member.
// The arithmetic is the members' own, so the results equal those of stepping
// each context by itself, bit for bit.
// ----------------------------------------------------------------------------
enum KernelId : int {
  KID_ROWS_D = 0,   // k_rows<double, NRB, CK>; variant = NRB * 100 + CK
  KID_ROWS_F,       // k_rows<float, NRB, 0>; variant = NRB
  KID_REDUCE,       // k_reduce
  KID_EPILOGUE,     // k_epilogue<CK>; variant = CK
  KID_JACOBI_WAVE,  // k_jacobi_wave (two decompositions per record)
  KID_JACOBI_WG,    // k_jacobi_svd
  KID_SHIFT,        // k_shift_sym
  KID_UNSHIFT,      // k_unshift
  KID_DE_PREP,      // k_de_prep<kLds>; variant = kLds
  KID_FLOW,         // k_spectral_flow
  KID_SSTEP,        // k_s_substep<kLds>; variant = kLds
  KID_SET_DIAG,     // k_set_diag
  KID_COPY,         // a device-to-device copy
  KID_COUNT
};

constexpr int kRecArgBytes = 448;

struct LaunchRec {
  int kid = 0, variant = 0;
  unsigned grid = 0, block = 0, lds = 0, nbytes = 0;
  alignas(16) unsigned char args[kRecArgBytes];
};

struct Recorder {
  std::vector<LaunchRec> recs;
};

inline Recorder*& current_recorder() {
  static thread_local Recorder* rec = nullptr;
  return rec;
}

// Append a launch to the installed recorder; false (nothing recorded) without one.
template <class A>
inline bool record_launch(int kid, int variant, unsigned grid, unsigned block, size_t lds, const A& a) {
  static_assert(sizeof(A) <= kRecArgBytes && std::is_trivially_copyable<A>::value, "recordable arguments");
  Recorder* R = current_recorder();
  if (!R) return false;
  R->recs.emplace_back();
  LaunchRec& L = R->recs.back();
  L.kid = kid;
  L.variant = variant;
  L.grid = grid;
  L.block = block;
  L.lds = (unsigned)lds;
  L.nbytes = sizeof(A);
  std::memcpy(L.args, &a, sizeof(A));
  return true;
}

// The batched forms: `count` members; start (device memory, 2 count + 1 ints)
// holds the workgroup offsets start[0..count], member i owning workgroups
// start[i] .. start[i+1] - 1 of the `total` launched, then each member's slot
// in `tab`: member i's argument block is tab + start[count + 1 + i] * kRecArgBytes.
hipError_t launch_rowops_batched(int kid, int variant, const void* tab, const int* start, int count, unsigned total,
                                 unsigned block, size_t lds, hipStream_t st);
hipError_t launch_de_batched(int kid, int variant, const void* tab, const int* start, int count, unsigned total,
                             unsigned block, size_t lds, hipStream_t st);

// The member of workgroup b, the last i with start[i] <= b.  A merged launch's
// table holds the workgroup offsets start[0..count], the members' argument
// slots start[count + 1 + i] and then the member of every workgroup, so this
// is one load (a binary search would be nine dependent ones per workgroup).
__device__ __forceinline__ int batch_member(const int* __restrict__ start, int count, int b) {
  return start[2 * count + 1 + b];
}

// r x r (up to 2r x 2r) FP64 matrices of one step, device-resident.
struct SmallState {
  double R1[kMaxRank * kMaxRank];      // first CholeskyQR factor
  double R1inv[kMaxRank * kMaxRank];
  double R2inv[kMaxRank * kMaxRank];   // second CholeskyQR factor, inverted
  double Rtot[kMaxRank * kMaxRank];    // R = R2 * R1
  double St[kMaxRank * kMaxRank];      // S~ of the KSL S sub-step
  double M[kMaxAug * kMaxRank];        // Galerkin U1' U0 (RA: 2r x r)
  double N[kMaxAug * kMaxRank];        // Galerkin V1' V0
  double Sinit[kMaxAug * kMaxAug];     // M S0 N'
  double Shat[kMaxAug * kMaxAug];      // RA core before the SVD
  double Ush[kMaxAug * kMaxAug];       // its left singular vectors
  double Vsh[kMaxAug * kMaxAug];       // its right singular vectors
  double sigma[kMaxAug];
  double Ctmp[kMaxAug * kMaxAug];      // scratch reduction result (EPI_STORE)
  double Cred[kMaxAug * kMaxAug];      // the reduced C handed from k_reduce to k_epilogue
  int r_new;
  int ndropped;
};

enum EpiKind {
  EPI_NONE = 0,
  EPI_CHOL1,   // C = Gram -> R1, R1^{-1}
  EPI_CHOL2,   // C = Gram -> R2^{-1}, Rtot = R2 R1; if S: S = Rtot (transpose: Rtot')
  EPI_STILDE,  // St = Rtot - C (transpose: St = (Rtot - C)')
  EPI_SFINAL,  // S = Rtot (transpose: Rtot')
  EPI_BUG_M,   // M = C
  EPI_BUG_N,   // N = C, Sinit = M S0 N'
  EPI_BUG_S,   // S = Sinit + C'
  EPI_RA_M,    // M = [I; C]
  EPI_RA_N,    // N = [I; C], Sinit = M S0 N'
  EPI_RA_S,    // Shat = Sinit + C'
  EPI_STORE,   // Ctmp = C
  EPI_COPY,    // dst = C (a x b; transpose: C')
  EPI_CHOL3,   // third factor of shifted CholeskyQR3: C = Gram -> R3^{-1} (into R2inv), Rtot = R3 Rtot; if S: S = Rtot
  EPI_COPYLTU, // dst = copyltu(X - C), X = S (ld ld_s) or 0, copyltu(Y) = tril(Y) + tril(Y, -1)' (a x a, ld a):
               // the thin-QR adjoint's coefficient
};

struct EpiArgs {
  int kind = EPI_NONE;
  int transpose = 0;
  int breakdown_is_error = 1;
  SmallState* small = nullptr;
  int* flags = nullptr;
  double* S = nullptr;          // state S (EPI_SFINAL / EPI_BUG_* / EPI_RA_N)
  int ld_s = 0;
  double* dst = nullptr;        // EPI_COPY
  int ld_dst = 0;               // EPI_COPY: leading dimension of dst (0: the rows of what is stored)
  // EPI_CHOL1: when > 0, factor G + s I with s = 11 (rows k + k (k + 1)) u trace(G)
  // (the shift of shifted CholeskyQR3 for an n x k factor with n = shift_rows)
  int64_t shift_rows = 0;
  // EPI_CHOL2 of a reduction out' [out | W] (b = 2a): also St = Rtot - R2^{-T} (out' W),
  // i.e. S~ = R - U1' W with U1 = out R2^{-1} (stored transposed when stilde == 2)
  int stilde = 0;
};

// Row kernel (see lri_rowops.hip).
template <typename PT>
struct RowArgs {
  int64_t rows = 0;
  int r_out = 0, r_in = 0;
  // sum of split partials part[s][c][i]
  const PT* parts = nullptr;
  int nsplit = 0, rb = 0;
  double* sum_out = nullptr;
  int64_t ld_sum = 0;
  // + base_sign * op(base), op = identity or the symmetric tridiagonal stencil
  // (st_d[i] on the diagonal, st_e[i] coupling rows i and i+1, st_e[rows-1]
  // coupling the last and first row when st_periodic)
  const double* base = nullptr;
  int64_t ld_base = 0;
  double base_sign = 1.0;
  const double* st_d = nullptr;
  const double* st_e = nullptr;
  int st_periodic = 0;
  // + xy_sign * X * op(Y)
  const double* X = nullptr;
  int64_t ldx = 0;
  const double* Y = nullptr;
  int ldy = 0, transY = 0;
  double xy_sign = 1.0;
  // outputs
  double* out = nullptr;
  int64_t ldo = 0;
  double* out2 = nullptr;
  int64_t ldo2 = 0;
  double* pk_f64 = nullptr;   // row-major [i][pk_cols]
  float* pk_f32 = nullptr;    // pk_layout 0: row-major [i][pk_cols]; 1: pass-T order (see pk_f32_offset)
  uint16_t* pk_hi = nullptr;  // fragment order of the BF16 passes (see pk_offset)
  uint16_t* pk_lo = nullptr;
  int pk_layout = 0;          // 0: pass-N order (BF16: 16-row blocks), 1: pass-T order (BF16: 64-row blocks)
  int pk_cols = 0;
  // reduction C = out' * W  (W == nullptr: out' * out; gram_w: out' * [out | W])
  int reduce = 0;
  const double* W = nullptr;
  int64_t ldw = 0;
  int rw = 0;
  int gram_w = 0;
  double* gpart = nullptr;    // [grid][r_out * (columns of C)]
  EpiArgs epi;
  int defer_epi = 0;          // reduce, but leave epi to the next row kernel's prologue
  // prologue: the deferred EPI_CHOL1/EPI_CHOL2 epilogue of the previous row
  // kernel on its reduced C (pre_a x pre_b), run by every workgroup (workgroup 0
  // stores its results); its R^{-1} (pre_a = r_in = r_out <= 32) then serves as Y
  EpiArgs pre;
  const double* pre_C = nullptr;
  int pre_a = 0, pre_b = 0;
  int max_blocks = 256;
  // filled by the launcher
  int tr = 0;  // rows per tile
};

template <typename PT>
hipError_t launch_rows(RowArgs<PT> A, hipStream_t st);

struct SvdArgs {
  const double* S = nullptr;  // k x k, ld lds
  int lds = 0;
  int k = 0;
  double* U = nullptr;  // k x k (ld k), columns by descending sigma
  double* V = nullptr;
  double* sigma = nullptr;
  int* r_out = nullptr;
  double tol = 0.0;
  int64_t rmax = 0;
  double* work = nullptr;  // global scratch of kSvdWorkDoubles for k > 64 (V, the QR preconditioner)
};
// the global Jacobi work of one decomposition with k > 64: V (ke x ke), the
// preconditioner's reflectors (k x k) and tau (k)
constexpr int kSvdWorkDoubles = 2 * kMaxAug * kMaxAug + kMaxAug;
hipError_t launch_jacobi_svd(SvdArgs P, hipStream_t st);

// The QR of a rank-deficient matrix.  CholeskyQR2 drops a pivot whose
// residual is at the rounding level and leaves a zero column in Q there; the
// unpivoted factorisation of the columns before it can then be ill-conditioned,
// so that Q R reproduces X only to ~eps cond.  Householder QR (qr! in the
// reference) returns an orthonormal Q and Q R = X to ~eps for any input.  When
// Rinv1(p, p) or Rinv2(p, p) of the CholeskyQR2 is zero for some column p,
// this recomputes the factorisation from X by classical Gram-Schmidt with
// reorthogonalisation (each column projected out twice): a column whose
// residual falls to the rounding level (|x| <= 16 k eps max_j |X(:, j)|) gets
// R(j, j) = 0 and, afterwards, an orthonormal completion q = e_i - Q Q(i, :)'
// for the row i of Q with the smallest norm, projected out twice and
// normalised.  Q (and its packed copy for the next pass) is rewritten, and the
// products formed from R: M = R - Q' W (mt: its transpose) and Z = R (zt: R').
// One workgroup; with nothing dropped it returns at once.
struct CompleteArgs {
  const double* X = nullptr;  // rows x k, ld rows (unchanged)
  double* Q = nullptr;        // rows x k, ld rows
  int64_t rows = 0;
  int k = 0;
  const double* Rinv1 = nullptr;  // k x k, ld k
  const double* Rinv2 = nullptr;
  double* pk_f64 = nullptr;
  float* pk_f32 = nullptr;
  uint16_t* pk_hi = nullptr;
  uint16_t* pk_lo = nullptr;
  int pk_layout = 0, pk_cols = 0;
  double* M = nullptr;
  int ldm = 0, mt = 0;
  const double* W = nullptr;
  int64_t ldw = 0;
  double* Z = nullptr;
  int ldz = 0, zt = 0;
};
hipError_t launch_complete_columns(const CompleteArgs& A, hipStream_t st);
// Reverse (adjoint) of the data-driven BUG step, r x r parts (one workgroup each).
struct BugAdjSmall {
  double RK[kMaxRank * kMaxRank], RL[kMaxRank * kMaxRank];          // R of the QRs of K1 and L1
  double RKinvT[kMaxRank * kMaxRank], RLinvT[kMaxRank * kMaxRank];  // their inverse transposes
  double Sb1[kMaxRank * kMaxRank];                                  // cotangent of S1
  double Mb[kMaxRank * kMaxRank], Nb[kMaxRank * kMaxRank];          // of M = U1'U0, N = V1'V0
  double Sb0[kMaxRank * kMaxRank];                                  // M' Sb1 N, then the full cotangent of S0
  double CK[kMaxRank * kMaxRank], CL[kMaxRank * kMaxRank];          // Kbar' U0, Lbar' V0
  double CqK[kMaxRank * kMaxRank], CqL[kMaxRank * kMaxRank];        // copyltu(-Ubar1' U1), copyltu(-Vbar1' V1)
  double St[kMaxRank * kMaxRank];                                   // KSL: S~ of the forward step
};
// Mb = Sb1 N S0', Nb = Sb1' M S0, Sb0 = M' Sb1 N, RK^{-T}, RL^{-T} (r x r, ld r; S0 ld lds)
hipError_t launch_bug_adj_pre(BugAdjSmall* A, const double* M, const double* N, const double* S0, int lds, int r,
                              hipStream_t st);
// Sb0 += CK' + CL
hipError_t launch_bug_adj_post(BugAdjSmall* A, int r, hipStream_t st);
// Reverse of the primal KSL step, r x r parts: stage 0: RK^{-T}, RL^{-T}, Mb = RL Sb1 (= R2 R2bar');
// stage 1: Nb = RK CL' (= R1 R1bar', R1bar = St bar = CL); stage 2: Sb0 = CK'.
hipError_t launch_ksl_adj_small(BugAdjSmall* A, int r, int stage, hipStream_t st);

// Reverse of the rank-adaptive BUG data step, small parts (k2 = 2r, r1 the new rank).
struct RaAdjSmall {
  double Ubs[kMaxAug * kMaxRank];    // Uhat' Ubar1 (k2 x r1, ld k2)
  double Vbs[kMaxAug * kMaxRank];    // Vhat' Vbar1
  double Sb1[kMaxRank * kMaxRank];   // cotangent of S1 (r1 x r1, ld r1; its diagonal enters)
  double Jm[kMaxAug * kMaxRank];     // Ush' Ubs, Vsh' Vbs (k2 x r1, ld k2)
  double Km[kMaxAug * kMaxRank];
  double core[kMaxAug * kMaxAug];    // the SVD rule's core, then core Vsh'
  double Shb[kMaxAug * kMaxAug];     // cotangent of Shat (k2 x k2, ld k2)
  double G1[kMaxRank * kMaxAug];     // S0 N' Shb' (r x k2, ld r)
  double G2[kMaxRank * kMaxAug];     // S0' M' Shb (r x k2, ld r)
  double G3[kMaxAug * kMaxRank];     // Shb N S0' (k2 x r, ld k2)
  double G4[kMaxAug * kMaxRank];     // Shb' M S0 (k2 x r, ld k2)
  double Sb0[kMaxRank * kMaxRank];   // M' Shb N, the cotangent of S0 (r x r, ld r)
  double RU[kMaxRank * kMaxRank], RV[kMaxRank * kMaxRank];          // R of the completions Q_U, Q_V
  double RUinvT[kMaxRank * kMaxRank], RVinvT[kMaxRank * kMaxRank];
  double CqU[kMaxRank * kMaxRank], CqV[kMaxRank * kMaxRank];        // copyltu(-Qbar' Q)
  double ZU[kMaxRank * kMaxRank], ZV[kMaxRank * kMaxRank];          // Zbar' U0, Zbar' V0
  double KU[kMaxRank * kMaxRank], LV[kMaxRank * kMaxRank];          // K1' U0, L1' V0
};
// stage 0: Jm, Km, RU^{-T}, RV^{-T};  stage 1: core;  stage 2: core <- core Vsh';  stage 3: Shb = Ush core,
// stage 4: G1..G4, Sb0 (SVD rule: Shb = Ush [(F o (J - J')) Sig + diag(sb) + Sig (F o (K - K'))] Vsh',
// F_ij = 1 / (sig_j^2 - sig_i^2) off the diagonal and 0 where i, j >= r1)
hipError_t launch_ra_adj_small(RaAdjSmall* A, const SmallState* st, const double* S0, int lds, int r, int r1,
                               int stage, hipStream_t s);

// C = A B' for k x k matrices (ld k), k <= kMaxAug, one workgroup.
hipError_t launch_small_abt(const double* A, const double* B, int k, double* C, hipStream_t st);
// two independent decompositions in one launch (one workgroup each)
hipError_t launch_jacobi_svd2(SvdArgs P0, SvdArgs P1, hipStream_t st);

// Product passes (see lri_tsmm.hip).
enum PassMode { PASS_F64 = 0, PASS_BF16X3 = 1, PASS_F32 = 2 };

struct PassArgs {
  const void* anext = nullptr;  // A[k_next], column-major, ld lda
  const void* aprev = nullptr;  // A[k_prev]; nullptr streams A[k_next] itself
  int dtype = 0;                // LRI_F64 / LRI_F32 / LRI_BF16
  int64_t lda = 0;
  int64_t n = 0, m = 0;
  int rb = 0;                   // output columns (multiple of 16 for F64, 32 otherwise)
  // B operand: F64 row-major [k][rb]; F32 row-major [k][rb]; BF16X3 hi/lo panels in fragment order
  const void* B = nullptr;
  const void* B2 = nullptr;
  int64_t ldb = 0;
  void* part = nullptr;         // [splits][rb][out_rows], double (F64) or float
  int splits = 0;
  int64_t chunk = 0;            // reduction-range length per split
};

// Choose splits/chunk for a pass whose output has `out_rows` rows and whose
// reduction runs over `red_len`; returns the partial buffer element count.
int64_t plan_pass(int mode, bool transposed, int64_t out_rows, int64_t red_len, int rb, int* splits,
                  int64_t* chunk);
// The same with a given reduction chunk (rounded up to the 256-row quantum).
int64_t plan_pass_chunk(int64_t out_rows, int64_t red_len, int rb, int64_t chunk_in, int* splits, int64_t* chunk);
// The chunks the autotuner times: those for about 128 .. 3072 workgroups, distinct.
int pass_chunk_candidates(int mode, bool transposed, int64_t out_rows, int64_t red_len, int rb, int64_t* chunks,
                          int cap);
hipError_t launch_passN(int mode, const PassArgs& a, hipStream_t st);
hipError_t launch_passT(int mode, const PassArgs& a, hipStream_t st);

hipError_t selftest_mfma(int which, const void* a, const void* b, void* out);
// Bare streaming-read bandwidth (TB/s) over two 16384^2 BF16 matrices; pattern 0 contiguous,
// 1: 8 columns x 128 B per wave instruction, 2: 1 column x 1 KB.
hipError_t stream_probe(int pattern, double* tbps);

// Matrix-ODE sub-steps (see lri_de.hip).  The forcing has rank q <= kMaxForce.
constexpr int kMaxForce = 4;

// y_j' = (T + mu_j I) y_j + sum_k g_k coef[k + j q], T symmetric tridiagonal
struct RK4ColArgs {
  int side = 0;               // host bookkeeping: 0 = the n-side operator A, 1 = the m-side operator B
  int64_t n = 0;
  int ncols = 0;
  const double* X = nullptr;  // initial columns, n x ncols (ld ldx)
  int64_t ldx = 0;
  double* Y = nullptr;        // result (may alias X)
  int64_t ldy = 0;
  const double* st_d = nullptr;
  const double* st_e = nullptr;
  int periodic = 0;
  const double* mu = nullptr;    // ncols
  const double* g = nullptr;     // n x q (ld ldg)
  int64_t ldg = 0;
  const double* coef = nullptr;  // q x ncols
  int q = 0;
  double h = 0.0;
  int64_t nsub = 0;
};
// One launch for the columns of two sub-problems (P1.ncols may be 0), e.g. the
// independent K and L steps of the unconventional integrators.
hipError_t launch_rk4_columns(const RK4ColArgs& P0, const RK4ColArgs& P1, hipStream_t st);

// Small preparation for a rotated sub-problem: CW = op(C) W (k x k), coef = FX W (q x k).
hipError_t launch_de_prep(const double* C, int ldc, int transC, const double* W, const double* FX, int k, int q,
                          double* CW, double* coef, hipStream_t st);

// S' = sgn (At S + S Bt + ga hb') in the eigenbases At = Wa diag(alpha) Wa', Bt = Wb diag(beta) Wb'
struct SStepArgs {
  int k = 0, l = 0, q = 0;
  const double* S0 = nullptr;
  int lds0 = 0;
  int transS0 = 0;              // start from S0'
  double* S1 = nullptr;
  int lds1 = 0;
  const double* Wa = nullptr;   // k x k
  const double* alpha = nullptr;
  const double* Wb = nullptr;   // l x l
  const double* beta = nullptr;
  const double* gaT = nullptr;  // q x k: (G' U)
  const double* hbT = nullptr;  // q x l: (H' V)
  double sgn = 1.0;
  double h = 0.0;
  int64_t nsub = 0;
  double* work = nullptr;       // 2 k l + (k + l) kMaxForce doubles
};
hipError_t launch_s_substep(const SStepArgs& P, hipStream_t st);

// Quadratic right-hand sides F(Y) = A Y + Y B' + G H' - (D Y) o Y (see lri_quad.hip).
constexpr int kQuadMaxRank = 32;  // columns of a sub-problem with the quadratic term

// part[g][c + k (a + k b)] = sum over the 64-row chunk g of X_ra Y_rb Y_rc,
// X = D Y (tridiagonal dl/dd/du, periodic dper) when dd != nullptr, else Y;
// launch_tensor3 then sums the chunks in order into out (k^3).
struct TensorArgs {
  int64_t rows = 0;
  int k = 0;
  const double* Y = nullptr;
  int64_t ldy = 0;
  const double *dl = nullptr, *dd = nullptr, *du = nullptr;
  int dper = 0;
  double* part = nullptr;  // ceil(rows / 64) * k^3
};
hipError_t launch_tensor3(const TensorArgs& Q, double* out, hipStream_t st);

// One RK4 stage (s = 0..3) of X' = A X + X M + g hv - N(X), N_ic = sum_ab (P X)_ia X_ib T_abc
// (P = D when dd != nullptr, else I), over all n rows; X0 holds the state of the
// RK4 step and is advanced in place by stage 3.
struct QuadStageArgs {
  int64_t n = 0;
  int k = 0;
  int64_t ld = 0;  // of X0, Xs, Xn, Acc
  double* X0 = nullptr;
  const double* Xs = nullptr;
  double* Xn = nullptr;
  double* Acc = nullptr;
  int stage = 0;
  double h = 0.0;
  const double* st_d = nullptr;  // A: symmetric tridiagonal (nullptr: none)
  const double* st_e = nullptr;
  int periodic = 0;
  const double* M = nullptr;     // k x k (nullptr: none)
  const double* g = nullptr;     // n x q
  int64_t ldg = 0;
  int q = 0;
  const double* hv = nullptr;    // q x k
  const double* T = nullptr;     // k^3 (nullptr: no quadratic term)
  const double *dl = nullptr, *dd = nullptr, *du = nullptr;
  int dper = 0;
};
hipError_t launch_quad_stage(const QuadStageArgs& Q, hipStream_t st);

// N RK4 steps of S' = sgn (Ga S + S Gb + (gaT' hbT) - P(S)) (k x l), one workgroup.
struct QuadSArgs {
  int k = 0, l = 0, q = 0;
  const double* S0 = nullptr;
  int lds0 = 0;
  int transS0 = 0;
  double* S1 = nullptr;
  int lds1 = 0;
  const double* Ga = nullptr;   // k x k
  const double* Gb = nullptr;   // l x l
  const double* gaT = nullptr;  // q x k
  const double* hbT = nullptr;  // q x l
  const double* W = nullptr;    // k^3, W[c + k (e + k f)] (nullptr: no quadratic term)
  const double* T = nullptr;    // l^3
  double sgn = 1.0;
  double h = 0.0;
  int64_t nsub = 0;
  double* work = nullptr;       // k l k + l l k doubles
};
hipError_t launch_quad_sstep(const QuadSArgs& Q, hipStream_t st);

// Reverse (adjoint) of the quadratic sub-steps (see lri_quad_adj.hip).
// One reverse RK4 stage s of X' = f(X) (f as in QuadStageArgs) over all rows:
// with Ybar the cotangent of the RK4 step's result and Xbp that of stage s + 1's
// input, kbar = (h w_s / 6) Ybar + c_s h Xbp (w = 1, 2, 2, 1; c = 1/2, 1/2, 1),
//   Xbo = J_f(X)' kbar,   Acc (=: s = 3 from Ybar, else +=) Xbo,
// and each workgroup adds its rows' parameter cotangents into its slice of
// part: [k^3: Tbar][k^2: Mbar][q k: hvbar][1: <kbar, A X>].
struct QuadAdjStageArgs {
  int64_t n = 0;
  int k = 0;
  int64_t ld = 0;
  const double* X = nullptr;
  const double* Ybar = nullptr;
  const double* Xbp = nullptr;  // nullptr for s = 3
  double* Xbo = nullptr;
  double* Acc = nullptr;
  int stage = 0;
  double h = 0.0;
  const double* st_d = nullptr;
  const double* st_e = nullptr;
  int periodic = 0;
  const double* M = nullptr;
  const double* g = nullptr;
  int64_t ldg = 0;
  int q = 0;
  const double* T = nullptr;
  const double *dl = nullptr, *dd = nullptr, *du = nullptr;
  int dper = 0;
  double* part = nullptr;
  int grid = 0;  // workgroups (filled by the launcher)
};
__host__ __device__ inline int quad_adj_part_len(int k, int q) { return k * k * k + k * k + q * k + 1; }
// workgroups of the reverse stage for `rows` rows (each owns one slice of part)
int quad_adj_grid(int64_t rows, int k);
hipError_t launch_quad_adj_stage(QuadAdjStageArgs Q, hipStream_t st);
// out[e] = sum over g < count of part[g * len + e] (in order)
hipError_t launch_part_sum(const double* part, int count, int len, double* out, hipStream_t st);

// Reverse of the S sub-step (N RK4 steps of S' = sgn (Ga S + S Gb + gaT' hbT - P(S))),
// one workgroup: recomputes the trajectory from S0, then returns the cotangent
// Sb0 of S0 (k x l, ld k) from that of the result (Sb1) and the cotangents of
// Ga, Gb, gaT, hbT, W, T (written, same layouts).
struct QuadSAdjArgs {
  int k = 0, l = 0, q = 0;
  const double* S0 = nullptr;
  int lds0 = 0;
  const double* Sb1 = nullptr;
  int ldsb1 = 0;
  double* Sb0 = nullptr;
  const double *Ga = nullptr, *Gb = nullptr, *gaT = nullptr, *hbT = nullptr, *W = nullptr, *T = nullptr;
  double sgn = 1.0;
  double h = 0.0;
  int64_t nsub = 0;
  double *Gab = nullptr, *Gbb = nullptr, *gaTb = nullptr, *hbTb = nullptr, *Wb = nullptr, *Tb = nullptr;
  double* work = nullptr;  // quad_sadj_work(k, l, nsub) doubles
};
inline int64_t quad_sadj_work(int k, int l, int64_t nsub) {
  const int64_t r = k > l ? k : l;
  return 4 * nsub * k * l + 3 * r * r * r;  // the stage inputs, then three r^3 intermediates
}
hipError_t launch_quad_sadj(const QuadSAdjArgs& Q, hipStream_t st);

// Cotangent of a basis X (rows x k) from those of its Galerkin data:
//   out = base + [tensor] + A X (Gb + Gb') + F fb + X2 op(Y2) + X3 op(Y3),
// tensor: the cotangent through W_abc = sum_i (D X)_ia X_ib X_ic (D = I when
// dd == nullptr) given Tb (k^3, [c + k (a + k b)]).  Every term is optional.
struct BasisVjpArgs {
  int64_t rows = 0;
  int k = 0;
  const double* X = nullptr;
  int64_t ldx = 0;
  const double *dl = nullptr, *dd = nullptr, *du = nullptr;
  int dper = 0;
  const double* Tb = nullptr;
  const double* st_d = nullptr;
  const double* st_e = nullptr;
  int periodic = 0;
  const double* Gb = nullptr;
  const double* F = nullptr;
  int64_t ldf = 0;
  int q = 0;
  const double* fb = nullptr;
  const double* X2 = nullptr;
  int64_t ldx2 = 0;
  const double* Y2 = nullptr;
  int ldy2 = 0, transY2 = 0;
  const double* X3 = nullptr;
  int64_t ldx3 = 0;
  const double* Y3 = nullptr;
  int ldy3 = 0, transY3 = 0;
  // optional second cotangents of the same Galerkin data (added to Tb, Gb, fb)
  const double* Tb2 = nullptr;
  const double* Gb2 = nullptr;
  const double* fb2 = nullptr;
  const double* base = nullptr;
  int64_t ldb = 0;
  double* out = nullptr;
  int64_t ldo = 0;
};
hipError_t launch_basis_vjp(const BasisVjpArgs& Q, hipStream_t st);
// out[0] = sumK[len-1] + <Gab0, Ga0> + <Gab1, Ga1>, out[1] = sumL[len-1] + <Gbb0, Gb0> + <Gbb1, Gb1>
// (the derivatives along the scales of the two linear parts; the first pair of
// each sum k x k, the second k1 x k1, ld k and k1)
hipError_t launch_quad_adj_scales(const double* sumK, const double* sumL, int len, const double* Gab0,
                                  const double* Ga0, const double* Gab1, const double* Ga1, const double* Gbb0,
                                  const double* Gb0, const double* Gbb1, const double* Gb1, int k, int k1,
                                  double* out, hipStream_t st);

// Analytic eigenbasis of a constant symmetric tridiagonal d I + e (S + S')
// (Dirichlet, or periodic with n >= 3): Q n x n column-major, lam n.
hipError_t launch_eig_tridiag_const(double* Q, double* lam, int64_t n, double d, double e, int periodic,
                                    hipStream_t st);
// Z (n x k, ld n) <- N RK4 steps of z_ij' = (lam_i + mu_j) z_ij + (fbar coef)_ij, in closed form.
hipError_t launch_spectral_flow(double* Z, int64_t n, int k, const double* lam, const double* mu, const double* fbar,
                                const double* coef, int q, double h, int64_t N, hipStream_t st);

// Symmetric eigen-decomposition M = W diag(ev) W' of a k x k matrix (k <= kMaxAug).
// work: k*k + 2 doubles; svd_work: kSvdWorkDoubles; Uscr: k*k doubles; r_scr: one int.
struct SymEigArgs {
  const double* M = nullptr;
  int ldm = 0, k = 0;
  double *W = nullptr, *ev = nullptr, *work = nullptr, *Uscr = nullptr, *svd_work = nullptr;
  int* r_scr = nullptr;
};
// one or two (count = 1, 2) independent decompositions
hipError_t launch_sym_eig(const SymEigArgs* P, int count, hipStream_t st);

// Element (row i, column c) of a BF16 hi/lo panel with rb columns, in the
// order the passes load their B fragments: pass N reads 8 consecutive rows
// for (16-row block, half h, column c); pass T reads 8 consecutive rows for
// (64-row block, k-step q, half h, column c).
__host__ __device__ inline int64_t pk_offset(int layout, int64_t i, int c, int rb) {
  if (layout == 0) {
    const int r = (int)(i & 15);
    return ((((i >> 4) * 2 + (r >> 3)) * rb) + c) * 8 + (r & 7);
  }
  const int r = (int)(i & 63);
  return (((((i >> 6) * 4 + ((r & 31) >> 3)) * 2 + (r >> 5)) * rb) + c) * 8 + (r & 7);
}

// Element (row i, column c) of an FP32 B panel with rb columns: pass N reads it
// row-major (layout 0); pass T (layout 1) reads, for the 32-row block of i, half h
// = (i / 16) % 2 and column c, the 16 consecutive rows 16 h .. 16 h + 15.
__host__ __device__ inline int64_t pk_f32_offset(int layout, int64_t i, int c, int rb) {
  if (layout == 0) return i * rb + c;
  return ((((i >> 5) * 2 + ((i >> 4) & 1)) * rb) + c) * 16 + (i & 15);
}

}  // namespace lri
