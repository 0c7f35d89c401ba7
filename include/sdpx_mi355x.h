/*
 * sdpx_mi355x.h -- C-ABI of the MI355X (gfx950) Newton-step provider for SDPX.
 *
 * One opaque handle owns the device copy of one core-form SDP
 *   min c'x  s.t.  X_l = sum_i x_i A_i^(l) - C_l >= 0,  l = 1..L   (n_eq = 0)
 * together with every device workspace buffer the dense Schur route needs,
 * and one HIP stream.  Each entry point is one phase of SDPX's newton_step!
 * (step.jl:643-1006) or of the KKT backend seam (kkt_backend.jl:211-530);
 * the host keeps the interior-point control logic (controller, restarts,
 * termination) and reads back only the scalars listed per call.
 *
 * Conventions
 *   - every call returns an int: 0 ok, 1 not positive definite, 2 non-finite,
 *     < 0 error (-1 invalid argument, -2 HIP error, -3 out of memory);
 *     sdpx_last_error() gives a message for the calling thread's last error;
 *   - host buffers are caller-owned and only read / written during the call;
 *   - matrices are column-major FP64; a block X_l is k_l x k_l and the blocks of
 *     a "concatenated" argument follow each other (offset sum_{l'<l} k_l'^2);
 *   - Av for block l is the k_l^2 x m panel of DenseCons.Av[l]
 *     (types/constraints.jl:26-28): column i is vec(A_i^(l));
 *   - only the lower triangle of the Schur matrix is authoritative
 *     (workspace.jl:1166-1172, docs/src/providers.md:45-50);
 *   - a handle is not thread-safe: one handle per SDPX Workspace, called
 *     sequentially (ext/SDPXMultiFloatLinearAlgebraExt.jl:14-20).
 */
#ifndef SDPX_MI355X_H
#define SDPX_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdpx_dense sdpx_dense_t;

#define SDPX_OK 0
#define SDPX_NOT_POSITIVE_DEFINITE 1
#define SDPX_NON_FINITE 2
#define SDPX_INVALID_ARGUMENT (-1)
#define SDPX_HIP_ERROR (-2)
#define SDPX_OUT_OF_MEMORY (-3)

/* ---- library ------------------------------------------------------------ */
const char* sdpx_last_error(void);
/* ABI version (major * 100 + minor). */
int sdpx_version(void);
/* Number of visible devices (0 when none). */
int sdpx_device_count(int* count);

/* ---- lifecycle: replaces instantiate_la_backend (la_backend.jl:508-559) and
 *      the Workspace constructor's buffer inventory (workspace.jl:779-1268) --- */
int sdpx_dense_create(int device, int L, const int* k, int m, sdpx_dense_t** out);
int sdpx_dense_destroy(sdpx_dense_t* h);
/* Device bytes held by the handle. */
int sdpx_dense_memory(sdpx_dense_t* h, long long* bytes);

/* ---- problem upload: DenseCons panels (ingest.jl:680-694, 950-962) -------- */
/* Av: k_l^2 x m column-major (bytes of DenseCons.Av[l]); C: k_l x k_l. */
int sdpx_dense_set_block(sdpx_dense_t* h, int l, const double* Av, const double* C);
int sdpx_dense_set_objective(sdpx_dense_t* h, const double* c);
/* Block l given as the reference's SparseBlockCOO (types/constraints.jl:59-109), as
 * stored (1-based): positions p = 1..na touch constraints ids[p] (here 0-based, strictly
 * increasing), entries ptr[p]..ptr[p+1]-1 with row, col, lin = (col-1)k + row and val,
 * every stored entry of the symmetric A (both triangles).  Replaces the block's dense
 * panel: the Schur build assembles its pairs as sparse_schur_block! does
 * (schur.jl:936-1032: W1 = Y A_p X^-1, S_pr = <W1, A_r>) and scatters them
 * (reduce_sparse_schur!, schur.jl:1229-1318); the Av passes read a densified copy.
 * Set once per block; with a sparse block, equilibration and the reduced-precision
 * Gram routes are rejected (the block assembles in FP64 as given). */
int sdpx_dense_set_block_sparse(sdpx_dense_t* h, int l, int na, const int* ids, const int* ptr, const int* lin,
                                const int* row, const int* col, const double* val, const double* C);

/* ---- iterate I/O (interior_point.jl:1528-1650 starts; Checkpoint problems.jl:27-38) */
int sdpx_dense_set_iterate(sdpx_dense_t* h, const double* x, const double* X, const double* Y);
int sdpx_dense_get_iterate(sdpx_dense_t* h, double* x, double* X, double* Y);
/* dx (m) and concatenated dX, dY blocks of the last direction. */
int sdpx_dense_get_direction(sdpx_dense_t* h, double* dx, double* dX, double* dY);
/* Schur matrix S (m x m); only the lower triangle is written by the build. */
int sdpx_dense_get_schur(sdpx_dense_t* h, double* S);
/* Overwrite S (test hook of kkt_regressions.jl:927-1000: copy_owned!(ws.S, .)). */
int sdpx_dense_set_schur(sdpx_dense_t* h, const double* S);
/* Overwrite the right-hand side r = -(d + v) (m) used by solve / refine. */
int sdpx_dense_set_rhs(sdpx_dense_t* h, const double* r);
int sdpx_dense_get_rhs(sdpx_dense_t* h, double* r);
int sdpx_dense_set_dx(sdpx_dense_t* h, const double* dx);

/* ---- a3: residuals + block factorization -----------------------------------
 * threaded_compute_residuals! (kernels/threaded.jl:423-505) / compute_residuals!
 * (step.jl:25-72) + factor_blocks! (step.jl:82-92).
 * mu: L centering targets; affine != 0 drops the mu*I term (use_affine).
 * out[0] = p_res, out[1] = d_res, out[2] = 1.0 when every block factor succeeded.
 * Returns 1 when a block Cholesky failed (factor != 0).                      */
int sdpx_dense_residuals(sdpx_dense_t* h, const double* mu, int affine, int factor, double* out);

/* ---- a4: dense Schur assembly S = sum_l Phat_l' Phat_l, lower triangle ------
 * threaded_schur_build! (threaded.jl:1298-1354) -> schur_build! (schur.jl:511-547). */
int sdpx_dense_schur_build(sdpx_dense_t* h);

/* ---- a6: Schur Cholesky with the sqrt(eps)*10^(a-1) ladder -----------------
 * factorize!(::DenseCholeskyBackend) (kkt_backend.jl:211-221) ->
 * _factor_dense_kkt_native! (kkt.jl:1404-1464) with la_chol! (la_backend.jl:1612-1622).
 * reg_attempts: 0..6. Returns 0 ok, 1 not positive definite after 6 attempts. */
int sdpx_dense_schur_factor(sdpx_dense_t* h, int* reg_attempts);

/* ---- a9: right-hand sides --------------------------------------------------
 * predictor: threaded_predictor_corrector_rhs! (threaded.jl:542-572) and
 * r = -(d + v) (step.jl:700-706).                                            */
int sdpx_dense_predictor_rhs(sdpx_dense_t* h);
/* corrector: mehrotra != 0 -> threaded_mehrotra_corrector_rhs! (threaded.jl:916-1061)
 * with sigma, mu_global and block_local; otherwise threaded_corrector_rhs!
 * (threaded.jl:837-888) with the L targets mu and beta. Then r = -(d + v)
 * (step.jl:863-865).                                                         */
int sdpx_dense_corrector_rhs(sdpx_dense_t* h, int mehrotra, double sigma, double mu_global, int block_local,
                             const double* mu, double beta);

/* ---- a7: direction solve dx = L^-T L^-1 r ----------------------------------
 * solve_direction! (kkt_backend.jl:392-410) -> _solve_dense_kkt_owned!
 * (kkt.jl:2570-2595) with la_trsv_lower! / la_trsv_transpose! (la_backend.jl:1646-1655). */
int sdpx_dense_solve(sdpx_dense_t* h);

/* ---- a8: refinement ----------------------------------------------------------
 * _kkt_direction_residual! (kkt.jl:3689-3708): rho = r - S dx (unfactored S,
 * schur_mul! kkt.jl:3514-3577). out[0] = ||rho||_inf, out[1] = ||r||_inf.     */
int sdpx_dense_direction_residual(sdpx_dense_t* h, double* out);
/* refine! (kkt_backend.jl:476-485) -> _refine_native_direction! (kkt.jl:3891-3963).
 * fixed_policy != 0: exactly max_steps passes with roll-back; otherwise stop at
 * abstol, on a decrease smaller than 2x, or after max_steps. */
int sdpx_dense_refine(sdpx_dense_t* h, double abstol, int max_steps, int fixed_policy, int* steps,
                      double* residual);

/* ---- a9: direction recovery dX = sum dx_i A_i + P, dY = sym(X^-1 (R - dX Y))
 * threaded_direction_blocks! (threaded.jl:574-644).                          */
int sdpx_dense_direction_blocks(sdpx_dense_t* h);

/* ---- a10: complementarity diagnostics (step.jl:373-544) -------------------
 * out[l] = <X_l, Y_l>, out[L + l] = <X_l + tX dX_l, Y_l + tY dY_l>.         */
int sdpx_dense_complementarity(sdpx_dense_t* h, double tX, double tY, double* out);

/* ---- a11: step length / PSD checks ----------------------------------------
 * fraction_to_boundary_search! (step.jl:1081-1098) on fraction_to_boundary_bound!
 * (kernels/generic.jl:350-420): t = fraction * bound when bound < 1.         */
int sdpx_dense_fraction_to_boundary(sdpx_dense_t* h, double primal_fraction, double dual_fraction, double* tX,
                                    double* tY);
/* line_search! (step.jl:1025-1075) with trial_has_cholesky_margin!
 * (kernels/generic.jl:318-340): first t = gamma^j (t *= gamma) accepted, or the
 * first t < min_step. */
int sdpx_dense_backtrack(sdpx_dense_t* h, double gamma, double min_step, double minimum_ratio, double* tX,
                         double* tY);

/* ---- a12: accepted-step update ---------------------------------------------
 * threaded_update_blocks! (threaded.jl:665-747) + x update (interior_point.jl:2228):
 * out[0] = sum_l <X_l, Y_l>, out[1] = 1.0 when X, Y, x are finite,
 * out[2 + l] = <X_l, Y_l>.                                                   */
int sdpx_dense_update(sdpx_dense_t* h, double tX, double tY, double* out);
/* _accepted_sdp_trial_residuals! (interior_point.jl:1287-1311):
 * d <- (1-tY) d - tY rho; out[0] = |1-tX| * block_primal_residual, out[1] = ||d||_inf. */
int sdpx_dense_accepted_residuals(sdpx_dense_t* h, double tX, double tY, double block_primal_residual,
                                  double* out);
/* pObj = c'x, dObj = sum_l <C_l, Y_l> (interior_point.jl:7-14, 2302-2303). */
int sdpx_dense_objectives(sdpx_dense_t* h, double* pobj, double* dobj);
/* solution_residuals (validation.jl:15-58): recomputed p_res, d_res. */
int sdpx_dense_solution_residuals(sdpx_dense_t* h, double* p_res, double* d_res);
/* restart rescale (interior_point.jl:2180-2195): which = 1 -> X, 2 -> Y. */
int sdpx_dense_scale_blocks(sdpx_dense_t* h, int which, double factor);
/* Cholesky diagonal qualities (step.jl:233-259, 1229): out[0] primal margin,
 * out[1] dual margin, out[2] Schur factor quality. */
int sdpx_dense_factor_quality(sdpx_dense_t* h, double* out);
/* Device-event timing of the Schur assembly and factor phases on the handle's
 * stream (adds one event synchronization per timed call). */
int sdpx_dense_set_profiling(sdpx_dense_t* h, int enabled);
/* out[0] Schur Gram kernel ms (sum), out[1] its launch count, out[2] panel
 * (L_X^-1 A_i M_Y) ms, out[3] split-K reduction ms, out[4] factor ms (all
 * attempts), out[5] factor calls; reset != 0 clears the counters. */
int sdpx_dense_kernel_times(sdpx_dense_t* h, int reset, double* out);
/* split-K factor used by the Schur Gram kernel of this handle. */
int sdpx_dense_schur_split(sdpx_dense_t* h, int* split);
/* Plan-time autotuner for the Schur build of this handle's shape (no
 * counterpart in the reference, whose BLAS calls take no tiling choices): runs
 * and times on the device each of the two P-hat panel products on 128x128,
 * 64x64, 128x64 or 64x128 tiles and the Schur Gram product at up to five
 * split-K factors and three tile-order super-block sides, and keeps the
 * fastest.  The products run on the handle's A_i with dense lower-triangular
 * stand-ins for L_X^-1 and M_Y (restored afterwards), so the operands carry
 * values like a solve's.  It also picks the bisection levels per batched trial round of
 * sdpx_dense_fraction_to_boundary (4..7; the result is the same for any
 * value).  Overwrites the panel, Schur and trial buffers (call before solving
 * or between solves).  out (optional, 20 doubles): tile shape code of each
 * panel product (0..3 in the order above), split, super-block side, panel ms,
 * Gram ms, create-time split, the 2 x 4 panel timings (product-major), the
 * levels per round, and the modelled search ms for 4..7 levels. */
int sdpx_dense_autotune(sdpx_dense_t* h, double* out);
/* Applies tiling choices without timing them (no counterpart in the
 * reference): panel shape codes as reported by sdpx_dense_autotune, Gram split
 * in [1, workspace capacity], tile-order super-block side >= 1, bisection
 * levels per trial round 1..7.  Lets a caller re-apply a saved autotune
 * result. */
int sdpx_dense_set_tiling(sdpx_dense_t* h, int panel0, int panel1, int split, int group, int ftb_depth);
/* Host-only (no device call; no counterpart in the reference): the work list
 * of the FP64 Schur Gram launch for an m x m Schur matrix, `group` x `group`
 * tile super-blocks and `split` K ranges, as the kernel reads it: for each
 * work item in launch order (XCD-remapped), out[2i] = split index and
 * out[2i+1] = ti << 16 | tj of its 128 x 128 lower tile; ceil(m/128) *
 * (ceil(m/128)+1) / 2 * split items.  Each XCD's run lists its full tiles
 * first and its diagonal tiles, short-last-row strips and empty items after. */
int sdpx_gram_work_list(int m, int group, int split, int* out);
/* Wait for the handle's stream. */
int sdpx_dense_synchronize(sdpx_dense_t* h);

/* ---- setup on the device -------------------------------------------------------
 * equilibrate(prob::SDPProblem) dense path (ingest.jl:1432-1521): Ruiz passes
 * per block, A_i <- E A_i E and C <- E C E with e = 1/sqrt(row infinity-norms
 * over every A_i and C) (ruiz_iters < 0: :auto = 2..8 passes, stop when
 * max |log rownorm| <= 1/20; otherwise exactly ruiz_iters passes), then the
 * per-variable scale s_i = max_l ||A_i^(l)||_inf (1 when zero), A_i /= s_i,
 * c <- (c ./ s) / objective_scale with objective_scale = ||c ./ s||_inf
 * (_equilibration_variable_scales, ingest.jl:1252-1270).  Scales the handle's
 * panels, C and c in place; E (sum k_l, per-block row scales), s (m),
 * objective_scale and passes (L) are returned (null skips).  A solve of the
 * scaled problem maps back by unequilibrate (ingest.jl:1523-1560):
 * x = x^ ./ s, X = E^-1 X^ E^-1, Y = E Y^ E objective_scale. */
int sdpx_dense_equilibrate(sdpx_dense_t* h, int ruiz_iters, double* E, double* s, double* objective_scale,
                           int* passes);
/* The handle's current C blocks (concatenated) and c (after equilibration). */
int sdpx_dense_get_data(sdpx_dense_t* h, double* C, double* c);
/* _kkt_cold_start_initialization (solver/interior_point.jl:699-1287), n_eq = 0:
 * X = Y = I, one Schur build (the Gram of the A_i) and one factor; S x =
 * [<A_i, C>] and S v = c solved with that factor and refined when the
 * normalized residual exceeds max(sqrt(eps), eps_primal, eps_dual); X = A(x) - C,
 * Y = A(v); the normalized residuals gated; every block lifted into the PD
 * interior (_cold_start_psd_shift!), the identity mass floor and the
 * cross-centering shifts applied (cold_start.jl:288-428).  Leaves (x, X, Y) as
 * the handle's iterate; mu[l] = beta <X_l, Y_l> / k_l.  out (16): ok, reason (0
 * ok, 1 identity_kkt_factorization, 2 kkt_cold_start_solve, 3
 * kkt_cold_start_residuals, 4 psd shift, 5 identity mass, 6 centering), Schur
 * regularization attempts, primal and dual normalized KKT residuals,
 * refinement steps, normalized p_res and d_res, largest primal and dual
 * shifts, primal and dual margins, kappa after / degree, kappa after the mass
 * floor / degree. */
int sdpx_dense_cold_start(sdpx_dense_t* h, double beta, double eps_primal, double eps_dual, double* mu, double* out);

/* ---- block-arrow elimination ----------------------------------------------------
 * Replaces factorize!(::ArrowBackend) / solve!(::ArrowBackend) (kkt_backend.jl:235-258,
 * 330-333 -> factor_arrow_kkt!, kkt.jl:2132-2349, and solve_arrow_kkt!, kkt.jl:2977-3096)
 * for a handle whose Schur matrix has the block-arrow structure of ArrowWorkspace
 * (workspace.jl:318-400): owner[i] = the block whose active set alone holds variable i
 * (a local variable), or -1 for a variable that touches several blocks (global).  After
 * this call sdpx_dense_schur_factor eliminates the local variables block by block (each
 * local block factored with its own sqrt(eps)*10^(a-1) ladder; W = D^-1 C through the
 * coupling rows), factors Sred = S_GG - C'D^-1 C with the ladder of
 * _factor_with_relative_regularization! (kkt.jl:1695-1716) and reports the sum of the
 * attempts; every solve with the factor (direction, refinement, cold start) takes the
 * arrow solve.  The unfactored S stays the refinement operator.  At most 128 local
 * variables per block, at least one local variable, no equality border, precision 0;
 * once per handle. */
int sdpx_dense_set_arrow(sdpx_dense_t* h, const int* owner);
/* Global variable count (-1 when the route is not set), local blocks, largest local block. */
int sdpx_dense_get_arrow(sdpx_dense_t* h, int* ng, int* nloc, int* max_q);

/* ---- equality-constrained KKT (n_eq > 0) ---------------------------------------
 * The KKT system  S dx - B dy = r,  B' dx = p  of the dense route with n
 * equality constraints (B: m x n column-major, uploaded once).
 * sdpx_dense_equality_factor replaces the n > 0 branch of
 * _factor_dense_kkt_native! (kkt.jl:1466-1560) after sdpx_dense_schur_factor:
 * B~ = L_S^-1 B (blocked substitution with the Schur factor), the
 * column normalization B^ = B~ D, D_jj = 1/||B~_j||_inf
 * (_normalize_equality_panel_columns!, kkt.jl:281-304), Q = B^'B^ on MFMA and
 * its Cholesky, accepted when min diag > tau max diag with
 * tau = max(max(m, n) eps, min(sqrt(eps), requested)) and no exactly
 * duplicated columns (_cholesky_has_numerical_rank, kkt.jl:76-176).
 * out (4): status (0 ok, 1 Gram factor failed, 2 numerically rank deficient),
 * min and max |diag| of the factor of Q, tau.  Returns 1 unless status 0.  The
 * reference's pivoted-QR fallback for rank-deficient B (kkt.jl:1517-1560) is
 * not provided: status 2 is returned instead. */
int sdpx_dense_set_equalities(sdpx_dense_t* h, int n, const double* B);
int sdpx_dense_equality_factor(sdpx_dense_t* h, double requested, double* out);
/* _solve_dense_kkt_owned! (kkt.jl:2570-2595) for n > 0 with the handle's r:
 * dx (left in the handle, sdpx_dense_get_direction) and dy (n, host). */
int sdpx_dense_solve_equality(sdpx_dense_t* h, const double* p, double* dy);
/* B~ D (m x n), Q (n x n, lower) and D (n) of the last equality factor. */
int sdpx_dense_get_equality(sdpx_dense_t* h, double* Bt, double* Q, double* scale);

/* Condition estimates from the last Schur factor L (S = L L'); which: 1 = [0], 2 = [1..3],
 * 3 = both (the others are left 0 / infinite); out (4):
 * [0] cond_1(S) as xPOCON estimates it (||S||_1 of the unfactored S times Higham's
 *     estimate of ||S^-1||_1 from triangular solves on the device);
 * [1] _triangular_condition_estimate's (mixed_precision_kkt.jl:371-392) xTRCON product
 *     1 / (rcond_1(L) rcond_inf(L)), an upper bound of cond_1(S);
 * [2], [3] rcond_1(L), rcond_inf(L).
 * Infinite (or 0) when an estimate is not positive and finite.  Dense factor only. */
int sdpx_dense_factor_condition(sdpx_dense_t* h, int which, double* out);

/* ---- a13: mixed precision ----------------------------------------------------
 * The reference's mixed-precision template (kernels/mixed_precision_kkt.jl:
 * low-precision factor, target-precision residuals and refinement,
 * _try_factor_mixed_kkt! :972-1111, _refine_mixed_direction! :1634-1713) applied
 * to the FP64 target: precision 1 assembles the Schur Gram product from an
 * FP32 copy of Phat on FP32 MFMA, precision 2 from a bf16 hi / lo split of
 * Phat as three bf16 MFMA products accumulated in FP32 (BF16x3); S stays FP64
 * (lower) in both; 0 restores FP64. */
int sdpx_dense_set_precision(sdpx_dense_t* h, int precision);
/* Precision of the Schur factor (the template's low-precision factor,
 * _factor_float64_preconditioner! in _try_factor_mixed_kkt!,
 * kernels/mixed_precision_kkt.jl:1010-1030, with FP32 under the FP64 target):
 * 1 factors S in FP32 -- panels and trailing updates on FP32 MFMA, the 64 x 64
 * diagonal blocks in FP64 and rounded -- and leaves the factor in FP64 storage
 * for the solves, the condition estimate and the refinement; 0 restores FP64.
 * The plain dense route only (no block-arrow elimination, no equality border). */
int sdpx_dense_set_factor_precision(sdpx_dense_t* h, int precision);
/* Precision of the P-hat products that feed the BF16x3 Schur assembly
 * (precision 2 of sdpx_dense_set_precision; the low-precision bulk assembly of
 * kernels/mixed_precision_kkt.jl:972-1111): 1 forms P-hat_i = L_X^-1 A_i M_Y on
 * FP32 MFMA from an FP32 copy of each dense block's Av (k^2 x m floats, kept
 * until the handle is destroyed, refreshed after set_block / equilibrate) and
 * writes the bf16 planes directly; 0 (default) forms P-hat in FP64 and splits
 * it.  Blocks whose side is not a multiple of 4 keep the FP64 products. */
int sdpx_dense_set_panel_precision(sdpx_dense_t* h, int precision);
/* rho = r - S dx with S applied matrix-free in FP64,
 * (S dx)_i = <A_i, X^-1 (sum_j dx_j A_j) Y> (the exact operator the FP32
 * assembly approximates); out[0] = ||rho||_inf, out[1] = ||r||_inf. */
int sdpx_dense_exact_residual(sdpx_dense_t* h, double* out);
/* Iterative refinement of dx against the matrix-free FP64 operator with the
 * current factor as preconditioner: stop at ||rho||_inf <= abstol, after
 * max_steps, or when a correction does not decrease the residual (rolled back). */
int sdpx_dense_refine_exact(sdpx_dense_t* h, double abstol, int max_steps, int* steps, double* residual);

/* ==== Batched route: many independent single-block SDPs, one launch per phase ==
 * No counterpart in the reference: SDPX solves one problem per Workspace
 * (sequential re-solves only, prepared.jl:391-447).  Each batch member is the
 * dense-Schur Newton method above (one PSD block of side k <= 64, m <= 2048
 * constraints, n_eq = 0).  Every phase is one launch over the members still
 * iterating; the per-member interior-point control runs on the host over
 * scalars read back once per phase, so members terminate independently. */
typedef struct sdpx_batch sdpx_batch_t;

typedef struct sdpx_batch_options {
  /* SolverOptions fields (types/workspaces.jl:101-228) */
  double beta, gamma, eps_gap, eps_primal, eps_dual, min_step, omega_step, stall_tolerance, refine_tol;
  double adaptive_sigma_max;
  int iter_max, max_restarts, max_centering, stall_iterations, restart;
  int refine_policy;      /* 0 :auto / :adaptive, 1 :fixed */
  int refine_steps, refine_max_steps;
  int adaptive;           /* parameter_strategy: 1 :adaptive, 0 :fixed */
  int step_rule;          /* 0 :auto, 1 :backtrack, 2 :fraction_to_boundary */
  /* 0: plain FP64; 1: mixed precision, the Schur Gram product assembled from
   * FP32 Phat panels on FP32 MFMA (FP64 S and factor) and every direction
   * refined against the matrix-free FP64 Schur operator, the template of
   * kernels/mixed_precision_kkt.jl:972-1111, 1475-1713 with FP64 as the target:
   * the predictor to 1e-8 max(||r||, 1), the corrector to refine_target
   * max(||r||, 1), at most mixed_max_refine steps each.  A member whose
   * reduced-precision Schur matrix does not factor, whose Cholesky pivot ratio
   * (max L_ii / min L_ii)^2 (a lower bound of cond_2(S)) exceeds cond_limit,
   * or whose refinement misses its target redoes the step in FP64 and keeps
   * FP64 (per member).  cond_limit <= 0 disables the pivot-ratio guard;
   * sdpx_batch_default_options sets 1e6, 1e-10 and 2 (a refinement round is
   * one launch over every member still refining, so the budget is small: a
   * member that needs more rounds continues in FP64). */
  int precision;
  double cond_limit, refine_target;
  /* differentiation checkpoint: the first iterate with <X, Y>/k <= adjoint_mu
   * is kept (0 disables; see sdpx_batch_restore_checkpoint) */
  double adjoint_mu;
  int mixed_max_refine;
} sdpx_batch_options_t;

/* Fill `o` with the reference defaults (SolverOptions{Float64}()) and tolerance
 * `tol` for the gap, primal and dual criteria (public/settings.jl:308-338). */
int sdpx_batch_default_options(sdpx_batch_options_t* o, double tol);
/* B members; member p has one PSD block of side k[p] (1..64) and m[p] (1..2048) constraints. */
int sdpx_batch_create(int device, int B, const int* k, const int* m, sdpx_batch_t** out);
int sdpx_batch_destroy(sdpx_batch_t* h);
int sdpx_batch_memory(sdpx_batch_t* h, long long* bytes);
/* Dense data, members concatenated: Av (k_p^2 x m_p column-major each, the
 * DenseCons panel, types/constraints.jl:26-28), C (k_p^2 each), c (m_p each). */
int sdpx_batch_set_dense(sdpx_batch_t* h, const double* Av, const double* C, const double* c);
/* Sparse data: constraint j of the batch (members concatenated, sum m_p
 * constraints) has entries ptr[j]..ptr[j+1]-1 of (row, col, val), 0-based,
 * row <= col; A_j is the symmetric matrix they define (duplicates add).  The
 * panels are densified on the device (SparseBlockCOO, types/constraints.jl:59-109). */
int sdpx_batch_set_coo(sdpx_batch_t* h, const long long* ptr, const int* row, const int* col, const double* val,
                       const double* C, const double* c);
/* The reference's SparseBlockCOO as stored (types/constraints.jl:59-109, built by
 * build_block_coo :79-109): per member, m_p + 1 Int32 positions ptr (1-based,
 * ptr[0] = 1) followed in the concatenation by the next member's; the
 * member's entries (lin, row, col Int32 1-based, val) follow the previous
 * member's.  Every stored entry of the symmetric A_i is given (both
 * triangles), lin = (col - 1) k + row.  Each lower entry must mirror an upper
 * entry of the same value (else SDPX_INVALID_ARGUMENT); the upper ones are
 * loaded as by sdpx_batch_set_coo. */
int sdpx_batch_set_block_coo(sdpx_batch_t* h, const int* ptr, const int* lin, const int* row, const int* col,
                             const double* val, const double* C, const double* c);
/* Host-only (no device call): one block's SparseBlockCOO (side k, m positions,
 * layout as above) converted to this ABI's COO (out_ptr m + 1 int64 0-based,
 * out_row <= out_col 0-based, out_val); *out_nnz = entries kept.  Null
 * outputs skip; the outputs need at most ptr[m] - 1 entries. */
int sdpx_block_coo_convert(int k, int m, const int* ptr, const int* lin, const int* row, const int* col,
                           const double* val, long long* out_ptr, int* out_row, int* out_col, double* out_val,
                           long long* out_nnz);
/* Start point, members concatenated: x (m_p), X and Y (k_p^2 each, column-major). */
int sdpx_batch_set_start(sdpx_batch_t* h, const double* x, const double* X, const double* Y);
/* Runs the interior-point loop (interior_point.jl:1323-2420) of every member to
 * termination; *iterations = number of batched Newton steps launched. */
int sdpx_batch_solve(sdpx_batch_t* h, const sdpx_batch_options_t* o, int* iterations);
/* Per member: status (0 Optimal, 1 IterLimit, 2 NumericalBreakdown, 3 Stalled,
 * 4 InsufficientPrecision, 5 MaxRestartsExceeded), iterations, objectives, and
 * the returned point (concatenated like set_start).  Null pointers skip. */
int sdpx_batch_get_result(sdpx_batch_t* h, int* status, int* iterations, double* pobj, double* dobj, double* x,
                          double* X, double* Y);
/* Iteration trace of one member's last solve: *n rows of SDPX_BATCH_TRACE_COLS
 * doubles (iteration, pobj, dobj, p_res, d_res of the step, tX, tY, sigma,
 * affine primal / dual steps, regularization attempts, refinement steps,
 * outcome 0 update / 1 centering / 2 restart / 3 termination); at most cap rows
 * are copied. */
#define SDPX_BATCH_TRACE_COLS 13
int sdpx_batch_get_trace(sdpx_batch_t* h, int member, double* rows, int cap, int* n);
/* ---- host-only hooks on the per-member controller (no device call) -----------
 * sdpx_batch_policy_select replaces select_iteration_parameters!
 * (adaptive_parameters.jl:511-548) over select_parameters (:188-194, 253-453):
 * a controller built from `o` (AdaptiveIPMController(opts), :455-476), with
 * nhist history rows of (sigma, unstable != 0) and, when preset_fallback != 0,
 * a controller already in fallback, selects the parameters for the
 * diagnostics `diag` (17 doubles in IterationDiagnostics order, :63-83:
 * iteration, primal_residual, dual_residual, relative_gap, mu, mu_aff,
 * affine_primal_step, affine_dual_step, previous_primal_step,
 * previous_dual_step, backtracking_count, regularization, refinement_count,
 * factorization_quality, achieved_residual_reduction, primal_psd_margin,
 * dual_psd_margin).  out (16): sigma, primal and dual fraction, backtracking
 * factor, refinement tolerance, refinement max count, minimum step, restart
 * scale, regularization, fallback, fallback reason (0 none, 1
 * nonfinite_diagnostics, 2 degraded_factorization, 3 repeated_instability, 4
 * previous_instability), the controller's fallback flag, beta, gamma after the
 * selection, and the policy's sigma_max, sigma_min.
 * sdpx_batch_policy_record replaces record_and_update! (:597-700) on a fresh
 * controller: rec (11) = iteration, complementarity before / after, primal and
 * dual residual, primal and dual step, primal and dual residual after,
 * backtracking count, refinement count; out (4) = achieved residual
 * reduction, unstable, the controller's fallback flag, instability count. */
int sdpx_batch_policy_select(const sdpx_batch_options_t* o, const double* diag, int nhist, const double* hist,
                             int preset_fallback, double* out);
int sdpx_batch_policy_record(const sdpx_batch_options_t* o, const double* rec, double* out);

/* ---- differentiation (no counterpart in the reference) -----------------------
 * For the user's dual-form model  max b'y  s.t.  C - sum_i y_i A_i >= 0  with
 * primal matrix X* (the core Y) and multipliers y* (the core x), and the loss
 * L = <G, X*> + h'y*, sdpx_batch_adjoint writes dL/dC (k_p^2 per member,
 * symmetric) and dL/db (m_p per member) at the current iterate of every
 * member, by implicit differentiation of the central-path conditions
 * (sum x_i A_i - X - C = 0, <A_i, Y> = c_i, XY = mu I; HKM linearisation):
 *   W = sym(X^-1 G Y),  lambda = S^-1 (h - [<A_i, W>]),  A_l = sum lambda_i A_i,
 *   dL/dC = -(sym(X^-1 A_l Y) + W),  dL/db = lambda,
 * with S the Schur matrix refactored at the iterate.  G, h, dC, db are
 * concatenated over members like set_start.  Returns 1 when a member could not
 * be factored (its outputs are NaN). */
int sdpx_batch_adjoint(sdpx_batch_t* h, const double* G, const double* hv, double* dC, double* db);
/* `iterations` Newton steps on the central-path equations XY = mu I (mu fixed
 * at <X, Y>/k on entry) with primal and dual feasibility: moves a solved
 * iterate onto the central path, where the adjoint above is the exact
 * derivative of the central-path point. */
int sdpx_batch_center(sdpx_batch_t* h, int iterations);
/* The same Newton steps with mu fixed at mu[p] per member (NaN: its current <X,Y>/k); mu null
 * is sdpx_batch_center.  Used to take the adjoint at two points of the central path. */
int sdpx_batch_center_at(sdpx_batch_t* h, int iterations, const double* mu);
/* Replace each member's iterate by its differentiation checkpoint (members
 * without one keep their iterate); *restored = members restored, mu[p] = the
 * checkpoint's <X, Y>/k (NaN without one).  The adjoint is taken at the
 * checkpoint re-centred by sdpx_batch_center: the final iterate of a solve is
 * too far from the central path, relative to its mu, for the linearisation. */
int sdpx_batch_restore_checkpoint(sdpx_batch_t* h, int* restored, double* mu);
/* Plan-time autotuner of the batched route (no counterpart in the reference): the
 * grouped Schur Gram launch timed at the start point (set_start) under each
 * visiting order of its tiles -- 0 the largest K = k^2 first, 1 member-major, 2
 * members dealt to the 8 XCDs by work so that a member's tiles share one L2 --
 * `repeats` launches each after a warm-up; the fastest order is kept.  out (4) =
 * chosen order, median milliseconds of orders 0, 1, 2.  The Schur matrices of
 * the start point are left in the handle. */
int sdpx_batch_autotune(sdpx_batch_t* h, int repeats, double* out);
/* Re-applies a saved order (0, 1 or 2). */
int sdpx_batch_set_gram_order(sdpx_batch_t* h, int mode);
/* Schur matrix of one member (m_p x m_p, lower triangle authoritative). */
int sdpx_batch_get_schur(sdpx_batch_t* h, int member, double* S);
/* Schur assembly of every member at the uploaded start point (parity hook of
 * schur_build!, schur.jl:511-547); returns 1 when a start block is not PD. */
int sdpx_batch_schur_at_start(sdpx_batch_t* h);
/* Assembly precision of sdpx_batch_schur_at_start and sdpx_batch_autotune: 0
 * FP64, 1 the FP32 Gram product of the mixed-precision solve (the precision of
 * a solve is its options' field).  Replaces the precision switch of the
 * reference's mixed-precision KKT template (kernels/mixed_precision_kkt.jl:58-120). */
int sdpx_batch_set_precision(sdpx_batch_t* h, int precision);
/* Per member (B ints): Newton steps of the last solve whose Schur matrix was
 * assembled in reduced precision (precision 1). */
int sdpx_batch_get_mixed_iterations(sdpx_batch_t* h, int* out);
/* Per member (B ints): why the last solve stopped, the reference's termination
 * messages as codes (interior_point.jl:1742-2217, step.jl:665-936): 1 optimal;
 * 2 Cholesky of X or Y failed (the iterate left the interior); 3 Schur
 * complement not positive definite after the regularization ladder; 4 final
 * direction residual above max(sqrt(eps), eps_p, eps_d) max(||r||, 1); 5 step
 * collapse near the solution; 6 step collapse with insufficient progress; 7
 * step collapse with no centering or restart left; 8 non-finite iterate after
 * the update; 9 iteration limit; 10 + the stagnation detector's reason (1 no
 * progress, 2 too slow, 3 precision floor). */
int sdpx_batch_get_reasons(sdpx_batch_t* h, int* out);

/* ==== a14: Lorentz (second-order) cone kernels ===============================
 * Many cones Q_q = {(t, u): t >= ||u||} of arbitrary dimensions resident in
 * HBM.  A handle owns SDPX_SOC_SLOTS coordinate vectors of sum_c q_c doubles
 * each (cone c's coordinates follow cone c-1's, head first; the caller decides
 * what a slot holds) plus the per-cone NT scalars eta, eta^2 and ok flags.
 * Each call is one bandwidth-bound launch over every cone, in the reference's
 * operation order (no floating-point contraction).  Replaces the per-cone
 * loops over src/soc_lorentz_kernels.jl and src/soc.jl kernels in the native
 * SOC workspace (soc_native.jl). */
#define SDPX_SOC_SLOTS 8
typedef struct sdpx_soc sdpx_soc_t;

int sdpx_soc_create(int device, int ncones, const int* dims, sdpx_soc_t** out);
int sdpx_soc_destroy(sdpx_soc_t* h);
int sdpx_soc_memory(sdpx_soc_t* h, long long* bytes);
/* Host <-> device copies of one slot (sum_c q_c doubles); device copy between slots. */
int sdpx_soc_set_vector(sdpx_soc_t* h, int slot, const double* v);
int sdpx_soc_get_vector(sdpx_soc_t* h, int slot, double* v);
int sdpx_soc_copy_vector(sdpx_soc_t* h, int src, int dst);
/* Per-cone eta, eta^2 and ok (1 when the last NT scaling succeeded) of the last
 * sdpx_soc_nt_scaling call; null pointers skip. */
int sdpx_soc_get_scaling(sdpx_soc_t* h, double* eta, double* eta2, int* ok);
/* _soc_nt_scaling! (soc_lorentz_kernels.jl:72-184) for every cone: NT scaling
 * point into slot w and lambda into slot lambda from slots s and z.  A cone
 * whose checks fail keeps its previous w / lambda and gets ok = 0, eta = 0;
 * *nfailed counts them.  Outputs must not alias inputs or each other. */
int sdpx_soc_nt_scaling(sdpx_soc_t* h, int s, int z, int w, int lambda, int* nfailed);
/* dst = op(src) per cone with the scaling point in slot w and the stored eta:
 * op 0 W (_soc_nt_apply_w!, soc_lorentz_kernels.jl:187-205), 1 W^-1
 * (_soc_nt_apply_winv!, :208-226), 2 (W'W)^-1 (_soc_nt_apply_hs_inverse!,
 * :229-248).  dst may alias src, not w. */
int sdpx_soc_apply(sdpx_soc_t* h, int op, int w, int src, int dst);
/* dst = left o right (_soc_jordan!, soc.jl:94-112); dst may alias either. */
int sdpx_soc_jordan(sdpx_soc_t* h, int left, int right, int dst);
/* left o dst = right (_soc_jordan_solve!, soc_lorentz_kernels.jl:464-485).
 * Returns 1 with *nbad cones when a divisor is not interior (those cones are
 * left untouched; the reference throws ArgumentError). */
int sdpx_soc_jordan_solve(sdpx_soc_t* h, int left, int right, int dst, int* nbad);
/* Jordan inverse (_soc_inverse!, soc.jl:126-138); returns 1 like jordan_solve. */
int sdpx_soc_inverse(sdpx_soc_t* h, int src, int dst, int* nbad);
/* _soc_fraction_to_boundary (soc.jl:148-214) per cone for s + t ds; *min_bound
 * = the minimum over cones (NaN propagates); per_cone (ncones) optional. */
int sdpx_soc_fraction_to_boundary(sdpx_soc_t* h, int s, int ds, double* min_bound, double* per_cone);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sdpx_soc_fraction_to_boundary on caller-owned device
 * buffers, with no handle and no allocation.  offsets: device, ncones + 1
 * int64 (offsets[0] = 0, cone c occupies [offsets[c], offsets[c+1]), every
 * cone at least 1 long); s, ds: device, offsets[ncones] doubles each;
 * min_bound: device, one double, receives the minimum over cones (NaN
 * propagates); per_cone: optional device array of ncones doubles for the
 * per-cone bounds.  Without per_cone the bounds go to a static device array
 * of SDPX_SOC_NOALLOC_MAX_CONES doubles (32 MiB per device, held until the
 * process ends); more cones are rejected with SDPX_INVALID_ARGUMENT.  The
 * launches are queued on `stream` (a hipStream_t; null = the default stream)
 * and the call returns without waiting.  Calls that overlap on one device
 * share that array and race. */
#define SDPX_SOC_NOALLOC_MAX_CONES (1 << 22)
int sdpx_soc_fraction_to_boundary_noalloc(int device, int ncones, const long long* offsets, const double* s,
                                          const double* ds, double* min_bound, double* per_cone, void* stream);
/* Per-cone margins v0 - ||v_tail|| (_soc_margin, soc_lorentz_kernels.jl:36)
 * and the number of strictly interior cones (_soc_is_interior, soc.jl:122). */
int sdpx_soc_margins(sdpx_soc_t* h, int slot, double* margins, int* ninterior);
/* _native_soc_strict_step (soc_native.jl:2705-2745) over all cones: the full
 * step when allowed and strictly interior, else the first of 0.99*bound
 * halved up to 24 times whose trial point is interior in every cone, else 0. */
int sdpx_soc_strict_step(sdpx_soc_t* h, int state, int direction, double bound, int allow_full_step,
                         double* step);
/* Fixed-trace Q3 kernels (every cone of dimension 3):
 * _soc_fixed_trace_hkm_metric! (soc_lorentz_kernels.jl:353-370) into dst;
 * cones outside the cone interior are counted and left untouched. */
int sdpx_soc_hkm_metric(sdpx_soc_t* h, int primal, int dual, int dst, int* nfailed);
/* mode 0: _soc_fixed_trace_hkm_rhs_coordinates (soc_lorentz_kernels.jl:372-418),
 * v = primal residual; mode 1: _soc_fixed_trace_hkm_recovery! (:420-461),
 * v = primal direction.  target = the centering target; include_affine != 0
 * subtracts the affine product of slots affine_primal / affine_dual. */
int sdpx_soc_hkm_contract(sdpx_soc_t* h, int mode, int primal, int dual, int v, int affine_primal,
                          int affine_dual, double target, int include_affine, int dst);
/* ---- LA provider seam, host-buffer operations (la_backend.jl:1529-1769) -------
 * The la_* methods of the provider's LA backend on caller-owned host arrays
 * (Float64, column-major, leading dimensions), computed on the device: each call
 * uploads its operands to the handle's scratch, runs the device kernels and
 * writes the result back before returning.  One handle per LA backend, called
 * sequentially.  Status 1 = not positive definite (la_chol! returning false). */
typedef struct sdpx_la sdpx_la_t;
int sdpx_la_create(int device, sdpx_la_t** out);
int sdpx_la_destroy(sdpx_la_t* h);
/* la_chol! (la_backend.jl:1612-1622): cholesky!(Symmetric(A, :L)) in place; the lower triangle
 * becomes L, the upper one is kept.  Returns 1 (A untouched) when A is not positive definite or
 * a lower entry of A or L is not finite. */
int sdpx_la_chol(sdpx_la_t* h, int n, double* A, long long lda);
/* la_trsm! (:1634): X <- L^-1 X, L lower n x n (non-unit), X n x nrhs. */
int sdpx_la_trsm(sdpx_la_t* h, int n, int nrhs, const double* L, long long ldl, double* X, long long ldx);
/* la_trsv_lower! / la_trsv_transpose! (:1642-1653): x <- L^-1 x (trans 0) or L^-T x (trans 1). */
int sdpx_la_trsv(sdpx_la_t* h, int n, const double* L, long long ldl, double* x, int trans);
/* la_syrk! (:1595-1604): BLAS syrk('L', 'T'), the lower triangle of S (n x n) <- alpha P'P +
 * beta S with P k x n; the upper triangle is kept. */
int sdpx_la_syrk(sdpx_la_t* h, int n, int k, double alpha, const double* P, long long ldp, double beta, double* S,
                 long long lds);
/* la_mul_owned! (:1556-1577): C <- alpha op(A) op(B) + beta C (op = transpose when ta / tb). */
int sdpx_la_gemm(sdpx_la_t* h, int ta, int tb, int m, int n, int k, double alpha, const double* A, long long lda,
                 const double* B, long long ldb, double beta, double* C, long long ldc);
/* la_axpby_owned! (:1659-1686): y <- alpha x + beta y. */
int sdpx_la_axpby(sdpx_la_t* h, long long n, double alpha, const double* x, double beta, double* y);
/* la_dot, la_norminf (:1529-1541); an empty vector gives 0. */
int sdpx_la_dot(sdpx_la_t* h, long long n, const double* x, const double* y, double* out);
int sdpx_la_norminf(sdpx_la_t* h, long long n, const double* x, double* out);

/* ---- SOC metric assembly: the dense branch of _native_soc_add_metric!
 * (soc_native.jl:1119-1218) over every cone at once.  A: (sum_c q_c) x n
 * column-major (cone c's coefficient rows A_c, q_c x n, at its coordinate
 * offset; n = the Hessian's variables), uploaded once per handle. */
int sdpx_soc_set_constraints(sdpx_soc_t* h, int n, const double* A);
/* H = H0 + sum_c A_c' (W_c'W_c)^-1 A_c with the NT scaling point in slot w and
 * the eta of the last sdpx_soc_nt_scaling (n x n, both triangles written; H0
 * null -> 0; H null -> kept on the device only).  Evaluated as Atil' Atil,
 * Atil_c = W_c^-1 A_c (W^-1 is symmetric, so W^-1 W^-1 = (W'W)^-1), on the
 * FP64 MFMA Gram kernel; the reference's two GEMMs M A_c and A_c'(M A_c) agree
 * with it to rounding.  Returns 1 with *nfailed cones when a cone's last NT
 * scaling failed (nothing assembled; the reference reports such a cone as
 * NumericalBreakdown). */
int sdpx_soc_metric(sdpx_soc_t* h, int w, const double* H0, double* H, int* nfailed);
/* Device-event timing of each following call's main kernel (one event sync per call). */
int sdpx_soc_set_profiling(sdpx_soc_t* h, int enabled);
int sdpx_soc_last_kernel_ms(sdpx_soc_t* h, double* ms);
int sdpx_soc_synchronize(sdpx_soc_t* h);

#ifdef __cplusplus
}
#endif

#endif /* SDPX_MI355X_H */
