/*
 * lri.h -- C ABI of liblri_hip.so, the MI355X (gfx950) engine for the per-step
 * hot path of dynamical low-rank approximation (DLRA) on data problems and on
 * matrix ODEs with a tridiagonal Lyapunov-type right-hand side.
 *
 * The reference (LowRankIntegrators.jl) is pure Julia; its "operator API" for
 * this path is multiple dispatch on (problem, algorithm).  Each entry point
 * below replaces one piece of it; the Julia file:line it stands in for is
 * given with every declaration.
 *
 * Conventions
 *  - Every function returns an int status: LRI_OK (0) or a negative LRI_E*.
 *    No C++ exception crosses the ABI.  lri_last_error() gives the message.
 *  - A context is single-threaded.  Step calls are asynchronous and ordered on
 *    the context's HIP stream; device-side conditions (a non-positive-definite
 *    Gram matrix, non-finite values) are latched on the device and returned by
 *    the next synchronising call (lri_synchronize, lri_get_state,
 *    lri_get_rank, lri_step_bug_ra).
 *  - Matrices crossing the ABI are column-major (Julia's layout) with an
 *    explicit leading dimension.  U is n x r, S is r x r, V is m x r, and
 *    Y = U * S * V'.  The state is held in FP64 on the device.  The arrays of
 *    lri_set_state / lri_get_state and of the adjoint entry points may be host
 *    or device memory (the copies infer the direction).
 *  - Snapshots are n x m column-major arrays A_k.  Steps take snapshot
 *    indices k_prev, k_next and use the increment dA = A[k_next] - A[k_prev],
 *    formed on the fly (it is never materialised).
 */
#ifndef LRI_H
#define LRI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lri_ctx lri_ctx;

/* status codes */
#define LRI_OK 0
#define LRI_EINVAL (-1)      /* bad argument (the reference raises AssertionError/DimensionMismatch) */
#define LRI_EHIP (-2)        /* a HIP runtime call failed */
#define LRI_EOOM (-3)        /* device allocation failed */
#define LRI_EBREAKDOWN (-4)  /* a Gram matrix in CholeskyQR2 was numerically singular */
#define LRI_ENONFINITE (-5)  /* NaN/Inf reached a Gram matrix (e.g. NaN in a snapshot) */
#define LRI_ENOTSUP (-6)     /* combination not supported by this build */

/* snapshot storage types */
#define LRI_F64 0
#define LRI_F32 1
#define LRI_BF16 2

/* arithmetic of the n x m x r products (everything else is always FP64) */
#define LRI_PREC_F64 0   /* FP64 products; F32/BF16 snapshots are decoded exactly */
#define LRI_PREC_MIXED 1 /* BF16 snapshots: BF16 MFMA with hi/lo split (3 products), FP32 accumulate;
                            F32 snapshots: FP32 MFMA; F64 snapshots: FP64 (as LRI_PREC_F64) */

/* projector-splitting orders (projector_splitting.jl:1-3) */
#define LRI_PRIMAL 0
#define LRI_DUAL 1
#define LRI_STRANG 2

/* limits of this build */
#define LRI_MAX_RANK 64 /* state rank capacity; the rank-adaptive step works on 2r <= 128 */

/* ---- library / context ------------------------------------------------- */

/* Version as 10000*major + 100*minor + patch. */
int lri_version(void);

/* Static text for a status code. */
const char* lri_status_string(int status);

/* Message of the last failure on ctx; with ctx == NULL, the last failure of
 * lri_create on the calling thread. Never NULL. */
const char* lri_last_error(const lri_ctx* ctx);

/* Create a context on HIP device `device` for n x m data and ranks up to
 * rcap (1 <= rcap <= LRI_MAX_RANK, rcap <= min(n, m)).  Allocates all work
 * buffers once (the reference allocates its caches in alg_cache,
 * projector_splitting.jl:87-105, unconventional.jl:86-107,
 * rank_adaptive_unconventional.jl:106-131, and again on every rank change,
 * :133-169; here rank changes reuse the same buffers). */
int lri_create(lri_ctx** out, int device, int64_t n, int64_t m, int64_t rcap);

/* Free the context and everything it owns (bound snapshots are borrowed and
 * are not freed).  NULL is accepted. */
int lri_destroy(lri_ctx* ctx);

/* Select the product arithmetic (LRI_PREC_F64 / LRI_PREC_MIXED). */
int lri_set_precision(lri_ctx* ctx, int precision);

/* ---- state (the SVDLikeRepresentation u of DLRIntegrator, primitives.jl:54-63) */

/* Copy Y = U*S*V' (host, column-major FP64) into the device state.  Replaces
 * `u = deepcopy(prob.u0)` in init (projector_splitting.jl:110,
 * unconventional.jl:112, rank_adaptive_unconventional.jl:97).
 * U must have orthonormal columns (as the reference assumes). */
int lri_set_state(lri_ctx* ctx, int64_t r, const double* U, int64_t ldu, const double* S,
                  int64_t lds, const double* V, int64_t ldv);

/* Synchronise, then copy the state to host (the `deepcopy(integrator.u)` of
 * update_sol!, primitives.jl:82-90).  Any pointer may be NULL to skip that
 * factor.  *r_out receives the current rank. */
int lri_get_state(lri_ctx* ctx, int64_t* r_out, double* U, int64_t ldu, double* S, int64_t lds,
                  double* V, int64_t ldv);

/* Current rank (synchronising). */
int lri_get_rank(lri_ctx* ctx, int64_t* r_out);

/* ---- data (update_data!, data_integrator.jl:22-28) --------------------- */

/* Allocate `count` library-owned snapshot slots of type dtype, column pitch n;
 * BF16 slots take n + 64 elements where n elements are a multiple of 4 KB (such
 * a pitch makes some pairs of BF16 snapshots stream markedly slower in the
 * passes, which read two at the same offsets at once; lri_snapshots_bind takes
 * the caller's pitch, and the same padding is recommended there). */
int lri_snapshots_reserve(lri_ctx* ctx, int64_t count, int dtype);

/* Copy one host snapshot (column-major, leading dimension ld >= n, same
 * dtype as reserved) into slot k.  Stream-ordered with the steps. */
int lri_snapshot_upload(lri_ctx* ctx, int64_t k, const void* host, int64_t ld);

/* Use caller-owned device arrays as snapshots (borrowed; the caller keeps
 * them alive).  ld >= n is the column stride in elements. */
int lri_snapshots_bind(lri_ctx* ctx, const void* const* dev_ptrs, int64_t count, int dtype,
                       int64_t ld);

/* ---- steps ------------------------------------------------------------- */

/* One projector-splitting (KSL) step with dA = A[k_next] - A[k_prev].
 * order LRI_PRIMAL: primal_LT_step! (projector_splitting.jl:117-152,191-196)
 * order LRI_DUAL:   dual_LT_step!   (projector_splitting.jl:154-189,198-203)
 * (LRI_STRANG needs a midpoint snapshot: use lri_step_ksl_strang.) */
int lri_step_ksl(lri_ctx* ctx, int order, int64_t k_prev, int64_t k_next);

/* Strang KSL step: primal half step on A[k_mid]-A[k_prev], then dual half
 * step on A[k_next]-A[k_mid] (projector_splitting.jl:205-211). */
int lri_step_ksl_strang(lri_ctx* ctx, int64_t k_prev, int64_t k_mid, int64_t k_next);

/* Unconventional (BUG) step (unconventional.jl:121-164). */
int lri_step_bug(lri_ctx* ctx, int64_t k_prev, int64_t k_next);

/* Rank-adaptive BUG step (rank_adaptive_unconventional.jl:171-233).  The new
 * rank min(rmax, truncate_to_tolerance(sigma, tol)) is written to *r_new
 * (synchronising).  When that rank exceeds the context capacity rcap the step
 * returns LRI_ENOTSUP and leaves the state at the step's start (the reference's
 * default rmax is 2^62, rank_adaptive_unconventional.jl:18-23: a rank is never
 * clipped to the capacity silently). */
int lri_step_bug_ra(lri_ctx* ctx, int64_t k_prev, int64_t k_next, double tol, int64_t rmax,
                    int64_t* r_new);

/* Reverse (adjoint) of the BUG data step (unconventional.jl:121-164): with the
 * context's state (U0, S0, V0) at the start of the step from A[k_prev] to
 * A[k_next], overwrite the cotangents (Ub, Sb, Vb) of the step's result
 * (U1, S1, V1) (host, column-major n x r, r x r, m x r) with those of
 * (U0, S0, V0).  The forward step is recomputed (up to U1, V1, their R
 * factors, M and N); the state is left unchanged.  Four passes over the
 * increment.  The reference has no adjoint (Julia AD would differentiate
 * unconventional_step!); a ChainRulesCore rrule calls this per step (a
 * checkpointed driver recomputes the states). */
int lri_step_bug_adjoint(lri_ctx* ctx, int64_t k_prev, int64_t k_next, double* Ub, int64_t ldub, double* Sb,
                         int64_t ldsb, double* Vb, int64_t ldvb);

/* The same for the primal KSL data step (projector_splitting.jl:129-152,
 * 191-196): both QR factors carry cotangents there (S1 = R2', S~ = R1 - U1'dA V0).
 * Four passes over the increment. */
int lri_step_ksl_adjoint(lri_ctx* ctx, int64_t k_prev, int64_t k_next, double* Ub, int64_t ldub, double* Sb,
                         int64_t ldsb, double* Vb, int64_t ldvb);

/* The same for the dual KSL data step (projector_splitting.jl:166-189).  The
 * dual step is the primal step of the transposed problem, dual(U, S, V; dA) =
 * (V1', S1'', U1') for primal(V, S', U; dA') = (U1', S1', V1'): the reverse step
 * is the primal one run on the context's transposed view (the same buffers,
 * the passes over dA exchanged).  Four passes over the increment. */
int lri_step_ksl_dual_adjoint(lri_ctx* ctx, int64_t k_prev, int64_t k_next, double* Ub, int64_t ldub, double* Sb,
                              int64_t ldsb, double* Vb, int64_t ldvb);

/* The same for the Strang KSL step (projector_splitting.jl:205-211) over
 * A[k_prev] -> A[k_mid] (primal half) -> A[k_next] (dual half), as
 * lri_step_ksl_strang: the primal half is recomputed from a device copy of the
 * state, the dual half reversed at its result, then the primal half reversed.
 * Ten passes over the two increments.  The state is left unchanged. */
int lri_step_ksl_strang_adjoint(lri_ctx* ctx, int64_t k_prev, int64_t k_mid, int64_t k_next, double* Ub,
                                int64_t ldub, double* Sb, int64_t ldsb, double* Vb, int64_t ldvb);

/* Reverse of the rank-adaptive BUG data step (rank_adaptive_unconventional.jl:182-233)
 * with the new rank r_new held fixed (the rank the forward step returned; the
 * recomputed step must reproduce it, else LRI_EINVAL): on entry the first
 * r_new columns of Ub (n rows) and Vb (m rows) and the r_new x r_new Sb hold the
 * cotangents of the result (U1, S1 = diag(sigma), V1) -- only Sb's diagonal
 * enters -- on exit the first r columns and the r x r Sb hold those of
 * (U0, S0, V0); the arrays hold max(r, r_new) columns.  The step is recomputed
 * up to the SVD of the 2r x 2r core (three passes), then the SVD rule (distinct
 * singular values), the augmented bases' QR rule and four more passes.  The
 * derivative is the one along orthonormal bases (the step assumes U0'U0 = I).
 * The reference has no adjoint (Julia AD would differentiate
 * rankadaptive_unconventional_step!). */
int lri_step_bug_ra_adjoint(lri_ctx* ctx, int64_t k_prev, int64_t k_next, double tol, int64_t rmax, int64_t r_new,
                            double* Ub, int64_t ldub, double* Sb, int64_t ldsb, double* Vb, int64_t ldvb);

/* Greedy step on an SVD-like state (greedy_integrator.jl:94-104,106-111)
 * with X = A[k_next] itself (update_data!, data_integrator.jl:22-28):
 * U1 = orth(X V0), V1 = orth(X' U0), S1 = U1' X V1.  Three passes over X. */
int lri_step_greedy(lri_ctx* ctx, int64_t k_next);

/* Greedy step on a TwoFactorRepresentation Y = U Z' (greedy_integrator.jl:84-92):
 * the state's V holds Z (m x r, not orthonormal; S is not used).
 * Z1 = X' U0 and U1 = Q P' for the SVD X Z1 = Q Sigma P' (the orthogonal
 * polar factor of X Z1), X = A[k_next].  Two passes over X. */
int lri_step_greedy_twofactor(lri_ctx* ctx, int64_t k_next);

/* Greedy step on a MatrixHybridProblem (primitives.jl:36-41,
 * greedy_integrator.jl:41-46,73-82): the two-factor state Y = U Z' with Z in the
 * state's V slot.  The Z integrator advances Z over [t, t + dt] with U held by
 * `substeps` classical RK4 steps of Z' = f(Z, U, t) = F(U Z')' U -- F the
 * operator set by lri_set_operator_lyapunov / lri_set_operator_quadratic, the
 * form of the reference's hybrid test (test/data_informed_approximation.jl:75) --
 * then X = A[k_next] and U1 = Q P' for the SVD X Z1 = Q Sigma P'.  The
 * reference's Z integrator is an adaptive Tsit5 ODE integrator. */
int lri_step_hybrid_greedy(lri_ctx* ctx, int64_t k_next, double dt, int64_t substeps);

/* ---- matrix ODE (MatrixDEProblem, primitives.jl:13-17) --------------------
 *
 * The reference evaluates a user closure F(Y, t) on low-rank arguments inside
 * its K/L/S right-hand sides (projector_splitting.jl:52-80,
 * unconventional.jl:51-79, rank_adaptive_unconventional.jl:59-86).  A closure
 * cannot run on the device, so the operator is given as a descriptor:
 *
 *     F(Y) = A*Y + Y*B' + G*H'
 *
 * with A (n x n) and B (m x m) symmetric tridiagonal -- diagonal a_diag[i],
 * a_off[i] coupling rows i and i+1, and when a_periodic also a_off[n-1]
 * coupling rows n-1 and 0 (periodic arrays need n >= 3) -- and a forcing G*H'
 * of rank q <= LRI_MAX_FORCE (G n x q, H m x q; q = 0 for none).  Arrays are
 * host FP64 and are copied; a_off has n-1 entries (n when periodic).
 * The sub-problems are integrated with `substeps` classical RK4 steps of size
 * dt/substeps (the reference's default is adaptive Tsit5).  The sub-problem
 * columns are limited to n, m <= LRI_DE_MAX_ROWS. */
#define LRI_MAX_FORCE 4
#define LRI_DE_MAX_ROWS 16384

int lri_set_operator_lyapunov(lri_ctx* ctx, const double* a_diag, const double* a_off, int a_periodic,
                              const double* b_diag, const double* b_off, int b_periodic, int64_t q,
                              const double* G, int64_t ldg, const double* H, int64_t ldh);

/* Adds the quadratic term -(D Y) o Y (o: elementwise product) to the operator
 * set by lri_set_operator_lyapunov (which clears it):
 *
 *     F(Y) = A Y + Y B' + G H' - (D Y) o Y,   D = tridiag(dl, dd, du) on the n rows,
 *
 * dl[i] multiplying Y[i-1, :], du[i] Y[i+1, :] (n entries each; with
 * `periodic` row 0 couples to row n-1 through dl[0] and row n-1 to row 0
 * through du[n-1], else those two are unused).  Viscous Burgers with one
 * parameter sample per column (test/data_agnostic_approximation.jl:16-33,
 * Delta*rho - (grad*rho) .* rho): A = nu Laplacian, D = central gradient.
 * The K/L/S sub-problems (projector_splitting.jl:52-80, unconventional.jl:51-79,
 * rank_adaptive_unconventional.jl:59-86) then carry the tensors
 * T = V x V x V and W = (D U) x U x U and are integrated by RK4 stage by stage;
 * rank <= LRI_QUAD_MAX_RANK (the rank-adaptive step: 2 r <= LRI_QUAD_MAX_RANK). */
#define LRI_QUAD_MAX_RANK 32
int lri_set_operator_quadratic(lri_ctx* ctx, const double* dl, const double* dd, const double* du, int periodic);

/* Keep the state in the analytic eigenbasis of a constant-coefficient operator
 * (on = 1) or leave it (on = 0).  Needs lri_set_operator_lyapunov with constant
 * coefficients on both sides (d I + e (shift + shift'), Dirichlet or periodic)
 * and no quadratic term, else LRI_EINVAL.  The device state then holds
 * Qa' U and Qb' V (Qa, Qb the orthogonal eigenbases of A and B) and the matrix-ODE
 * steps use the operator's diagonal form, so a step makes no n x n transforms:
 * the DLRA steps commute with an orthogonal change of basis on either side (the
 * QRs, the Galerkin products and the SVD are basis-covariant), so the results
 * are those of the untransformed steps up to rounding.  lri_set_state /
 * lri_get_state / lri_trajectory_get transform at the boundary; a data-path step,
 * an adjoint or a new operator leaves the eigenbasis first.  No counterpart in
 * the reference (it evaluates F on the factors in the original basis,
 * projector_splitting.jl:52-80, unconventional.jl:51-79). */
int lri_set_resident_eigenbasis(lri_ctx* ctx, int on);

/* One KSL step of the matrix ODE over [t, t+dt]: order LRI_PRIMAL
 * (projector_splitting.jl:129-152,191-196), LRI_DUAL (:166-189,198-203) or
 * LRI_STRANG (primal over dt/2, then dual over dt/2, :205-211; each half with
 * `substeps` RK4 steps). */
int lri_step_de_ksl(lri_ctx* ctx, int order, double dt, int64_t substeps);

/* One BUG step of the matrix ODE (unconventional.jl:133-164). */
int lri_step_de_bug(lri_ctx* ctx, double dt, int64_t substeps);

/* One rank-adaptive BUG step of the matrix ODE (rank_adaptive_unconventional.jl:171-233);
 * truncation, *r_new and LRI_ENOTSUP above the capacity as lri_step_bug_ra
 * (synchronising). */
int lri_step_de_bug_ra(lri_ctx* ctx, double dt, int64_t substeps, double tol, int64_t rmax, int64_t* r_new);

/* Reverse (adjoint) of lri_step_de_bug (unconventional.jl:133-164 with the
 * sub-problem right-hand sides of :51-79): with the context's state
 * (U0, S0, V0) at the start of the step, overwrite the cotangents (Ub, Sb, Vb)
 * of the step's result (host, column-major n x r, r x r, m x r) with those of
 * (U0, S0, V0), and add to scale_bar[0], scale_bar[1] (if not NULL) the
 * derivatives along scales a, b of the two linear parts,
 * F = a A Y + b Y B' + G H' - (D Y) o Y at a = b = 1 (A = nu Laplacian:
 * dL/dnu = scale_bar[0] / nu).  The forward step is recomputed (keeping the RK4
 * sub-step starts of the K and L sub-problems); the state is left unchanged.
 * Needs the quadratic operator (lri_set_operator_quadratic; D = 0 for a linear
 * right-hand side).  The derivative is that of the discrete step: the same RK4
 * sub-steps and the QR with a positive diagonal of R.  The reference has no
 * adjoint (Julia AD would differentiate unconventional_step!); a ChainRulesCore
 * rrule calls this per step. */
int lri_step_de_bug_adjoint(lri_ctx* ctx, double dt, int64_t substeps, double* Ub, int64_t ldub, double* Sb,
                            int64_t ldsb, double* Vb, int64_t ldvb, double* scale_bar);

/* The same for the primal KSL matrix-ODE step (lri_step_de_ksl with LRI_PRIMAL;
 * projector_splitting.jl:52-80,129-152): the S sub-step runs backwards in time
 * from R1, and both QR factors carry cotangents (S1 = R2'). */
int lri_step_de_ksl_adjoint(lri_ctx* ctx, double dt, int64_t substeps, double* Ub, int64_t ldub, double* Sb,
                            int64_t ldsb, double* Vb, int64_t ldvb, double* scale_bar);

/* Reverse of lri_step_de_bug_ra (rank_adaptive_unconventional.jl:194-233 with the
 * sub-problem right-hand sides of :59-86) with the new rank r_new held fixed (the
 * rank the forward step returned; the recomputed step must reproduce it, else
 * LRI_EINVAL): on entry the first r_new columns of Ub (n rows) and Vb (m rows)
 * and the r_new x r_new Sb hold the cotangents of the result (U1, S1 = diag(sigma),
 * V1) -- only Sb's diagonal enters -- on exit the first r columns and the r x r
 * Sb hold those of (U0, S0, V0); the arrays hold max(r, r_new) columns.
 * scale_bar as for lri_step_de_bug_adjoint.  The step is recomputed up to the
 * SVD of the 2r x 2r core (the K and L sub-step starts kept), then the SVD rule
 * (distinct singular values), the RK4 reverse of the 2r S sub-step, the
 * augmented bases' QR rule and the K and L reverses.  Needs the quadratic
 * operator and 2 r <= LRI_QUAD_MAX_RANK.  The state is left unchanged. */
int lri_step_de_bug_ra_adjoint(lri_ctx* ctx, double dt, int64_t substeps, double tol, int64_t rmax, int64_t r_new,
                               double* Ub, int64_t ldub, double* Sb, int64_t ldsb, double* Vb, int64_t ldvb,
                               double* scale_bar);

/* ---- batches of independent problems -------------------------------------
 *
 * Many small solves of uneven n, m and rank (the reference runs them one
 * `solve` at a time, primitives.jl:68-80).  A batch owns one context per
 * problem.  Members are set up through their contexts (lri_batch_context:
 * state, operator); they are owned by the batch.
 *
 * Merged launches (the default): the members share one stream, and a
 * rank-adaptive batch step whose members are all in their resident eigenbasis
 * (lri_set_resident_eigenbasis) runs every stage of the step as one launch
 * over all members: each member's launches are recorded once per rank and step
 * parameters, their argument blocks kept in a device table, and position p of
 * all members' sequences launched together, each workgroup finding its member
 * in the table.  Each member's arithmetic is that of stepping its context
 * alone, so the results are the same bit for bit.  The merged list runs as a
 * captured HIP graph while it stays the same (lri_batch_use_graphs).  Other
 * batch steps run the members one after another on the shared stream.
 * Without merged launches (lri_batch_use_merged_launches(batch, 0)) each
 * member has its own stream and a batch step queues every member's step before
 * it waits for any. */
typedef struct lri_batch lri_batch;
int lri_batch_create(lri_batch** out, int device, int64_t count, const int64_t* n, const int64_t* m,
                     const int64_t* rcap);
int lri_batch_destroy(lri_batch* batch);
/* The message of the last failure ("problem i: ..."), or of a failed create. */
const char* lri_batch_last_error(const lri_batch* batch);
int lri_batch_context(lri_batch* batch, int64_t i, lri_ctx** ctx);
/* lri_step_de_bug on every member (substeps[i] RK4 sub-steps for member i). */
int lri_batch_step_de_bug(lri_batch* batch, double dt, const int64_t* substeps);
/* lri_step_de_bug_ra on every member; r_new[i] receives member i's new rank.
 * All members' steps are queued up to their SVD before the first rank is read.
 * Member i truncates at min(rmax, rcap[i]): its capacity is its own rmax.  With
 * merged launches the new ranks and condition flags of all members are read in
 * one transfer, a member whose step raised a condition keeps its state, and
 * the rotations to the new ranks are left queued (lri_batch_synchronize waits). */
int lri_batch_step_de_bug_ra(lri_batch* batch, double dt, const int64_t* substeps, double tol, int64_t rmax,
                             int64_t* r_new);
/* HIP graphs for rank-adaptive batch steps (default on = 1): with merged
 * launches, the merged list is captured once it repeats and replayed while it
 * stays the same; without, each member's step up to its SVD is a graph
 * captured for the member's current rank.  0: direct launches. */
int lri_batch_use_graphs(lri_batch* batch, int on);
/* Merged launches on (1, the default) or off (0); see above.  Switching waits
 * for every member's queued work. */
int lri_batch_use_merged_launches(lri_batch* batch, int on);
/* The number of batch steps that ran as merged launches, and the kernel
 * launches of the last of them (both phases and the gather of the new ranks). */
int lri_batch_launch_stats(const lri_batch* batch, int64_t* merged_steps, int64_t* launches);
/* The status of member i from the last batch call (LRI_OK or the LRI_E* that
 * member reported).  A batch step runs every member even when one fails: a
 * member whose step failed keeps its state at the step's start (a rank-adaptive
 * member that needs a rank above its capacity reports LRI_ENOTSUP), every other
 * member has advanced. */
int lri_batch_member_status(const lri_batch* batch, int64_t i, int* status);
int lri_batch_synchronize(lri_batch* batch);

/* ---- trajectory (DLRSolution.Y, update_sol!, primitives.jl:46-49,82-90) ----
 *
 * The reference deep-copies u into the solution after every step.  Here the
 * copies stay in device memory until they are read: lri_trajectory_save queues
 * a device-to-device copy of the current state into a slot (stream-ordered, no
 * synchronisation), lri_trajectory_get synchronises and copies one slot out
 * (host or device arrays, as lri_get_state).  Each slot holds
 * (n + m) * rcap + rcap^2 doubles. */
int lri_trajectory_reserve(lri_ctx* ctx, int64_t slots);
int lri_trajectory_save(lri_ctx* ctx, int64_t slot);
int lri_trajectory_get(lri_ctx* ctx, int64_t slot, int64_t* r_out, double* U, int64_t ldu, double* S, int64_t lds,
                       double* V, int64_t ldv);

/* Slots slot0 .. slot0+count-1 in one copy (synchronising), in their device
 * layout: slot i at host + i * (n*rcap + rcap*rcap + m*rcap) holds U (n x rcap,
 * ld n), S (rcap x rcap, ld rcap) and V (m x rcap, ld m), of which the first
 * ranks[i] columns (and the leading ranks[i] x ranks[i] block of S) are the
 * state.  Slots saved in the resident eigenbasis are rejected (LRI_EINVAL; use
 * lri_trajectory_get).  One transfer instead of three synchronising copies per
 * state, for DLRSolution.Y after a solve with save_everystep. */
int lri_trajectory_get_range(lri_ctx* ctx, int64_t slot0, int64_t count, int64_t* ranks, double* host);

/* Wait for all queued work; returns a latched device-side status, if any. */
int lri_synchronize(lri_ctx* ctx);

/* The context's HIP stream (hipStream_t), for callers that order their own
 * work (e.g. snapshot generation) with the steps. */
int lri_get_stream(lri_ctx* ctx, void** stream);

/* ---- measurement ------------------------------------------------------- */

/* Kernel families timed by lri_kernel_timing_read. */
#define LRI_KERNEL_PASS_N 0 /* dA * B, reduction over the m columns */
#define LRI_KERNEL_PASS_T 1 /* dA' * B, reduction over the n rows */
#define LRI_KERNEL_SMALL 2  /* all FP64 n x r / r x r kernels of a step */
#define LRI_KERNEL_DE 3     /* matrix-ODE sub-steps: eigenbasis transforms of the K/L columns
                               (FP64 passes over the n x n basis), or RK4 column integration */
#define LRI_KERNEL_COUNT 4

/* HIP-event timing on the context's stream: mode 0 disables it, mode 1 times
 * every launch of the families above, mode 2 only the streaming kernels
 * (LRI_KERNEL_PASS_N/T and LRI_KERNEL_DE, two events per launch, the least
 * intrusive).  Enabling resets the totals. */
int lri_kernel_timing(lri_ctx* ctx, int mode);

/* Synchronise and return the summed duration (ms) and launch count of a
 * kernel family since timing was enabled. */
int lri_kernel_timing_read(lri_ctx* ctx, int family, double* total_ms, int64_t* launches);

/* Time only one launch in `every` of each family (default 1: all), so that the
 * events disturb a timed loop less; kept across lri_kernel_timing calls. */
int lri_kernel_timing_sample(lri_ctx* ctx, int every);

/* Synchronise and copy up to cap per-launch durations (ms, in launch order) of
 * a family since timing was enabled; *count receives how many there are. */
int lri_kernel_timing_samples(lri_ctx* ctx, int family, float* ms, int64_t cap, int64_t* count);

/* ---- autotuning ----------------------------------------------------------
 *
 * The streaming passes split their reduction range into chunks, one partial
 * result per chunk; the default chunk aims at about 512 workgroups.  This
 * times every candidate chunk (those for about 128 .. 3072 workgroups) of
 * pass N and pass T over dA = A[kn] - A[kp] (kp < 0: A[kn]) at the context's
 * current rank and precision, in three interleaved rounds of `reps` launches
 * each (after an untimed one), and keeps the candidate with the lowest median
 * for this context's passes of that shape; the default chunk stays unless the
 * best is at least 1 % faster (the partial buffers grow if a candidate needs
 * it).  report (8 doubles): for pass N then pass T, the chosen chunk, its split
 * count, its median time per launch (us) and that of the default chunk (us).  A chosen chunk changes the order in which
 * the partials are summed, so results move within the precision's rounding.
 * No counterpart in the reference (its products are BLAS calls). */
int lri_autotune_passes(lri_ctx* ctx, int64_t kp, int64_t kn, int64_t reps, double* report);

/* ---- diagnostics --------------------------------------------------------- */

/* Run one matrix-core instruction on device 0 with the caller's operands and
 * return its raw per-lane results, so a test can check the operand/result
 * lane layouts the kernels assume.  which: 0 = FP64 16x16x4 (in: 64 + 64
 * doubles, out: 256 doubles), 1 = BF16 32x32x16 (in: 512 + 512 bf16 bit
 * patterns, out: 1024 floats), 2 = FP32 32x32x2 (in: 64 + 64 floats, out:
 * 1024 floats).  Results are ordered [lane][register]. */
int lri_selftest_mfma(int which, const void* a, const void* b, void* out);

/* Bandwidth context for the passes (no counterpart in the reference): what
 * bare non-temporal 16-byte loads stream, in TB/s, over two column-major
 * 16384 x 16384 BF16 matrices (1 GiB) on the current device, no arithmetic:
 * pattern 0 contiguous, 1 = a wave instruction covering 8 columns x 128 B (the
 * pass over dA' x U), 2 = 1 column x 1 KB.  Allocates and frees 1 GiB. */
int lri_stream_probe(int pattern, double* tbps);

/* One product pass with the current state and precision, reduced to FP64 and
 * copied to host (synchronising): which = 0 gives dA * V (n x r, ld ldo >= n),
 * which = 1 gives dA' * U (m x r, ld ldo >= m), dA = A[k_next] - A[k_prev].
 * These are the n x m x r products of the data sub-integrator
 * (data_integrator.jl:14-16) that every step is built from. */
int lri_debug_product(lri_ctx* ctx, int which, int64_t k_prev, int64_t k_next, double* out, int64_t ldo);

/* ---- allocation-free product ----------------------------------------------
 *
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the product pass of lri_debug_product without a context,
 * on caller-owned device arrays and asynchronously on `stream` (a hipStream_t;
 * NULL: the default stream), allocating nothing.
 *
 *   which = 0:  out (n x r, ld ldo) = dA * B,   B m x r (ld ldb)
 *   which = 1:  out (m x r, ld ldo) = dA' * B,  B n x r (ld ldb)
 *
 * with dA = A_next - A_prev, both n x m column-major (ld lda) of type dtype,
 * in the arithmetic `precision` selects (as lri_set_precision).  B and out are
 * FP64.  Every intermediate (the packed B operand and the split partials of
 * the pass) lives in a static device buffer of LRI_NOALLOC_SCRATCH_BYTES that
 * each GPU the library runs on holds until the process ends; a shape that
 * needs more (e.g. n = m = 65536 at r = 64 in FP64) returns LRI_EINVAL.  All
 * calls on one GPU share that buffer: calls that overlap on one GPU (two
 * streams, two host threads) race.  lri_noalloc_last_error() gives the message
 * of the calling thread's last failure. */
#define LRI_NOALLOC_SCRATCH_BYTES (32 * 1024 * 1024)
int lri_product_noalloc(int which, int dtype, int precision, int64_t n, int64_t m, const void* a_prev,
                        const void* a_next, int64_t lda, const double* B, int64_t ldb, int64_t r, double* out,
                        int64_t ldo, void* stream);
const char* lri_noalloc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LRI_H */
