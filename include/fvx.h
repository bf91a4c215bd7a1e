/* finitevolx MI355X (gfx950) hot-path library: the C ABI of libfvx_hip.so.
 *
 * Conventions
 *  - Fields are row-major [ny][nx] device arrays (float or double; suffix
 *    _f32 / _f64).  Operators that mirror the reference's ghost-ring contract
 *    take the FULL array (one ghost ring included) and write a zero ring.
 *  - Masks are device uint8 arrays of the same shape (0 = dry); NULL = all wet.
 *  - Every function takes a hipStream_t (as void*) and only enqueues work,
 *    except where noted (plan creation, and solves that report convergence,
 *    which synchronise the stream once per solve).
 *  - Return value: 0 = success, -1 = invalid argument, -2 = HIP error;
 *    fvx_last_error() returns the thread's last message.  Non-convergence of
 *    a solver is not an error (reported in the info structs).
 *  - Reference citations are into finitevolX 0.0.40 (paths under finitevolx/_src/).
 */
#ifndef FVX_H
#define FVX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* fvx_last_error(void);
int fvx_abi_version(void);
int fvx_device_synchronize(void);

/* ---- C-grid point stencils ------------------------------------------------
 * op codes (FVX_OP_*) and the reference method each one replaces:            */
enum {
  FVX_OP_DIFF_X_T_TO_U = 0,  /* Difference2D.diff_x_T_to_U  operators/difference.py:194 */
  FVX_OP_DIFF_Y_T_TO_V = 1,  /* Difference2D.diff_y_T_to_V  operators/difference.py:215 */
  FVX_OP_DIFF_Y_U_TO_X = 2,  /* Difference2D.diff_y_U_to_X  operators/difference.py:236 */
  FVX_OP_DIFF_X_V_TO_X = 3,  /* Difference2D.diff_x_V_to_X  operators/difference.py:258 */
  FVX_OP_DIFF_Y_X_TO_U = 4,  /* Difference2D.diff_y_X_to_U  operators/difference.py:280 */
  FVX_OP_DIFF_X_X_TO_V = 5,  /* Difference2D.diff_x_X_to_V  operators/difference.py:301 */
  FVX_OP_DIFF_X_U_TO_T = 6,  /* Difference2D.diff_x_U_to_T  operators/difference.py:326 */
  FVX_OP_DIFF_Y_V_TO_T = 7,  /* Difference2D.diff_y_V_to_T  operators/difference.py:347 */
  FVX_OP_DIVERGENCE = 8,     /* Difference2D.divergence / Divergence2D.__call__ :372, divergence.py:113 */
  FVX_OP_DIV_NOFLUX = 9,     /* Divergence2D.noflux         operators/divergence.py:138 */
  FVX_OP_CURL = 10,          /* Difference2D.curl           operators/difference.py:401 */
  FVX_OP_LAPLACIAN = 11,     /* Difference2D.laplacian      operators/difference.py:431 */
  FVX_OP_GRAD_PERP_U = 12,   /* Difference2D.grad_perp (u)  operators/difference.py:456 */
  FVX_OP_GRAD_PERP_V = 13,   /* Difference2D.grad_perp (v)  operators/difference.py:456 */
  FVX_OP_T_TO_X = 14,        /* Interpolation2D.T_to_X      operators/interpolation.py:158 */
  FVX_OP_V_TO_T = 15         /* Interpolation2D.V_to_T      operators/interpolation.py:206 */
};
/* out = op(a[, b]) * mask, zero ghost ring.  mask: the output stagger's mask. */
int fvx_stencil2d_f32(int op, const float* a, const float* b, float* out, int ny, int nx,
                      double dx, double dy, const uint8_t* mask, void* stream);
int fvx_stencil2d_f64(int op, const double* a, const double* b, double* out, int ny, int nx,
                      double dx, double dy, const uint8_t* mask, void* stream);

/* ---- Advection2D.__call__ (advection/advection.py:329-420) -----------------
 * method: FVX_ADV_*; limiter used when method == FVX_ADV_TVD.
 * tier_x / tier_y: NULL for the unmasked path, else per-cell largest usable
 * stencil in {0,2,4,6} (Mask2D.get_adaptive_masks, mask/cartesian.py:495-553)
 * for the mask-dispatchable methods (weno3, weno5, wenoz5, TVD;
 * advection/flux.py:100-253).  mask_h (nullable) multiplies the tendency.   */
enum {
  FVX_ADV_NAIVE = 0, FVX_ADV_UPWIND1 = 1, FVX_ADV_UPWIND2 = 2, FVX_ADV_UPWIND3 = 3,
  FVX_ADV_WENO3 = 4, FVX_ADV_WENOZ3 = 5, FVX_ADV_WENO5 = 6, FVX_ADV_WENOZ5 = 7, FVX_ADV_TVD = 8
};
enum { FVX_LIM_MINMOD = 0, FVX_LIM_VAN_LEER = 1, FVX_LIM_SUPERBEE = 2, FVX_LIM_MC = 3 };
int fvx_advection2d_f32(int method, int limiter, const float* h, const float* u, const float* v,
                        float* out, const uint8_t* tier_x, const uint8_t* tier_y,
                        const uint8_t* mask_h, int ny, int nx, double dx, double dy,
                        void* stream);
int fvx_advection2d_f64(int method, int limiter, const double* h, const double* u,
                        const double* v, double* out, const uint8_t* tier_x,
                        const uint8_t* tier_y, const uint8_t* mask_h, int ny, int nx, double dx,
                        double dy, void* stream);

/* ---- One rk3_ssp_step (timestepping/explicit_rk.py:74-121) of a doubly
 * periodic tracer with Advection2D.__call__ tendencies (advection.py:329-420):
 * h (ny, nx; ghost width ng) advances in place; its ghosts are filled
 * periodically (boundary/boundary.py:52-78) before each tendency, as the
 * reference's periodic advection tests do (tests/test_advection.py:558-628),
 * and on exit.  Method / limiter / tiers / mask as fvx_advection2d_*.  The two
 * stage arrays live in a per-device buffer the library allocates on first
 * use and grows when a call needs more.                                      */
int fvx_advection_rk3_step_f32(int method, int limiter, float* h, const float* u,
                               const float* v, const uint8_t* tier_x, const uint8_t* tier_y,
                               const uint8_t* mask_h, int ny, int nx, int ng, double dx,
                               double dy, double dt, void* stream);
int fvx_advection_rk3_step_f64(int method, int limiter, double* h, const double* u,
                               const double* v, const uint8_t* tier_x, const uint8_t* tier_y,
                               const uint8_t* mask_h, int ny, int nx, int ng, double dx,
                               double dy, double dt, void* stream);
/* The same step without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The two
 * stage arrays live in a 32 MiB static device array that each GPU holds until
 * the process ends; 2 * ny * nx * sizeof(element) above 32 MiB is refused with
 * -1 before any device work.  Calls on one GPU share that array: calls that
 * overlap on one GPU (other streams, other threads) race.                   */
int fvx_advection_rk3_step_noalloc_f32(int method, int limiter, float* h, const float* u,
                                       const float* v, const uint8_t* tier_x,
                                       const uint8_t* tier_y, const uint8_t* mask_h, int ny,
                                       int nx, int ng, double dx, double dy, double dt,
                                       void* stream);
int fvx_advection_rk3_step_noalloc_f64(int method, int limiter, double* h, const double* u,
                                       const double* v, const uint8_t* tier_x,
                                       const uint8_t* tier_y, const uint8_t* mask_h, int ny,
                                       int nx, int ng, double dx, double dy, double dt,
                                       void* stream);

/* ---- One Runge-Kutta stage of a tracer with an Advection2D tendency ---------
 * out = a hn + b hs + cdt L(hs), L the Advection2D.__call__ tendency
 * (advection.py:329-420; method / limiter / tiers / mask as fvx_advection2d_*)
 * fused with the stage combination of rk3_ssp_step
 * (timestepping/explicit_rk.py:74-121), the form the reference evaluates as
 * a tendency call plus array arithmetic.  out may alias hn, not hs.          */
int fvx_advection_stage_f32(int method, int limiter, const float* hs, const float* u,
                            const float* v, const float* hn, float* out, const uint8_t* tier_x,
                            const uint8_t* tier_y, const uint8_t* mask_h, int ny, int nx,
                            double dx, double dy, double a, double b, double cdt, void* stream);
int fvx_advection_stage_f64(int method, int limiter, const double* hs, const double* u,
                            const double* v, const double* hn, double* out,
                            const uint8_t* tier_x, const uint8_t* tier_y, const uint8_t* mask_h,
                            int ny, int nx, double dx, double dy, double a, double b,
                            double cdt, void* stream);

/* ---- Reconstruction2D face fluxes (advection/reconstruction.py:572-1307) ---
 * <method>_x (dir 0, east faces) / <method>_y (dir 1, north faces): every face
 * 1..n-2 along the direction on the interior of the other, zero ring; method /
 * limiter codes as above (FVX_ADV_WENOZ3 included).  tier (nullable): the
 * direction's stencil tiers for the *_masked variants (weno5, wenoz5, TVD;
 * also weno3), selected by the upwind cell (flux.py:100-253).                */
int fvx_face_flux_f32(int method, int limiter, int dir, const float* h, const float* vel,
                      float* out, const uint8_t* tier, int ny, int nx, void* stream);
int fvx_face_flux_f64(int method, int limiter, int dir, const double* h, const double* vel,
                      double* out, const uint8_t* tier, int ny, int nx, void* stream);

/* ---- Diffusion2D (diffusion/diffusion.py:120-329) --------------------------
 * kappa: device field or NULL (then kappa0 is used everywhere).             */
int fvx_diffusion2d_f32(const float* h, const float* kappa, double kappa0, float* out, int ny,
                        int nx, double dx, double dy, const uint8_t* mask_h,
                        const uint8_t* mask_u, const uint8_t* mask_v, void* stream);
int fvx_diffusion2d_f64(const double* h, const double* kappa, double kappa0, double* out,
                        int ny, int nx, double dx, double dy, const uint8_t* mask_h,
                        const uint8_t* mask_u, const uint8_t* mask_v, void* stream);
/* Diffusion2D.fluxes (diffusion/diffusion.py:290-329) */
int fvx_diffusion2d_fluxes_f32(const float* h, const float* kappa, double kappa0, float* fx,
                               float* fy, int ny, int nx, double dx, double dy,
                               const uint8_t* mask_u, const uint8_t* mask_v, void* stream);
int fvx_diffusion2d_fluxes_f64(const double* h, const double* kappa, double kappa0, double* fx,
                               double* fy, int ny, int nx, double dx, double dy,
                               const uint8_t* mask_u, const uint8_t* mask_v, void* stream);
/* BiharmonicDiffusion2D.__call__ (diffusion/diffusion.py:543-574), fused. */
int fvx_biharmonic2d_f32(const float* h, float* out, int ny, int nx, double dx, double dy,
                         double kappa, const uint8_t* mask_h, void* stream);
int fvx_biharmonic2d_f64(const double* h, double* out, int ny, int nx, double dx, double dy,
                         double kappa, const uint8_t* mask_h, void* stream);

/* ---- MomentumAdvection2D.__call__ (diffusion/momentum.py:37-248) ------------
 * (du, dv) = (+(zeta v)_u - dK/dx, -(zeta u)_v - dK/dy) on [2:-2, 2:-2], zero
 * elsewhere; scheme 0 energy (Sadourny E), 1 enstrophy (Sadourny Z), 2
 * Arakawa-Lamb (1/3 E + 2/3 Z).  Masks: all four (h, u, v, xy_corner_strict
 * stagger masks) or all NULL.                                               */
int fvx_momentum_advection2d_f32(const float* u, const float* v, float* du, float* dv, int scheme,
                                 const uint8_t* mask_h, const uint8_t* mask_u,
                                 const uint8_t* mask_v, const uint8_t* mask_xy, int ny, int nx,
                                 double dx, double dy, void* stream);
int fvx_momentum_advection2d_f64(const double* u, const double* v, double* du, double* dv,
                                 int scheme, const uint8_t* mask_h, const uint8_t* mask_u,
                                 const uint8_t* mask_v, const uint8_t* mask_xy, int ny, int nx,
                                 double dx, double dy, void* stream);

/* ---- arakawa_jacobian (operators/jacobian.py:22-131) -----------------------
 * f, g: (batch, ny, nx); out: (batch, ny-2, nx-2).                          */
int fvx_arakawa_jacobian_f32(const float* f, const float* g, float* out, int batch, int ny,
                             int nx, double dx, double dy, void* stream);
int fvx_arakawa_jacobian_f64(const double* f, const double* g, double* out, int batch, int ny,
                             int nx, double dx, double dy, void* stream);

/* ---- masked_laplacian (solvers/iterative.py:144-191); mask: same dtype ---- */
int fvx_masked_laplacian_f32(const float* psi, const float* mask, float* out, int ny, int nx,
                             double dx, double dy, double lambda_, void* stream);
int fvx_masked_laplacian_f64(const double* psi, const double* mask, double* out, int ny,
                             int nx, double dx, double dy, double lambda_, void* stream);

/* ---- enforce_periodic (boundary/boundary.py:52-78), ghost width ng, in place */
int fvx_enforce_periodic_f32(float* field, int ny, int nx, int ng, void* stream);
int fvx_enforce_periodic_f64(double* field, int ny, int nx, int ng, void* stream);

/* ---- rk3_ssp_step stage combination (timestepping/explicit_rk.py:104-121) --
 * out = a*x + b*y + c*z (+ d*w) over n elements (y, z, w nullable); out may
 * be one of the inputs (each element is read before it is written).         */
int fvx_lincomb3_f32(int64_t n, double a, const float* x, double b, const float* y, double c,
                     const float* z, float* out, void* stream);
int fvx_lincomb3_f64(int64_t n, double a, const double* x, double b, const double* y, double c,
                     const double* z, double* out, void* stream);
int fvx_lincomb4_f32(int64_t n, double a, const float* x, double b, const float* y, double c,
                     const float* z, double d, const float* w, float* out, void* stream);
int fvx_lincomb4_f64(int64_t n, double a, const double* x, double b, const double* y, double c,
                     const double* z, double d, const double* w, double* out, void* stream);

/* ---- layer_to_mode / mode_to_layer (vertical/vertical_modes.py:193-263) ---
 * out[k, :] = sum_l C[k, l] in[l, :]; C is a HOST row-major nl x nl matrix
 * (Cl2m or Cm2l from decompose_vertical_modes), nl <= 8, n points per layer. */
int fvx_layer_mix_f32(const float* in, float* out, const double* C, int nl, int64_t n,
                      void* stream);
int fvx_layer_mix_f64(const double* in, double* out, const double* C, int nl, int64_t n,
                      void* stream);

/* ---- Multigrid (solvers/multigrid.py) --------------------------------------
 * fvx_mg_plan_create replaces build_multigrid_solver (multigrid.py:905-1062):
 * mask / coeff are HOST float64 arrays (ny, nx) (coeff nullable = 1);
 * n_levels <= 0 = automatic (multigrid.py:50-76).  kind: 0 = the reference's
 * no-flux variable-coefficient operator (_apply_operator, :327-405);
 * 1 = masked_laplacian's Dirichlet operator with zero padding beyond the
 * array.  precision: 32, 64, or 3264 (FP32 hierarchy + FP64 fine operator for
 * mixed-precision solves).  Non-divisible grids fail with -1.             */
int fvx_mg_plan_create(const double* mask, const double* coeff, int ny, int nx, double dx,
                       double dy, double lambda_, int n_levels, int n_pre, int n_post,
                       int n_coarse, double omega, int kind, int precision, void** plan_out);
int fvx_mg_plan_destroy(void* plan);
/* Coarse-level wall correction of a Dirichlet plan's FP32 hierarchy (on = 1):
 * on level l each dry or outside neighbour takes the ghost value g_l u with
 * g_l = (1 - 2^l) / (1 + 2^l), which keeps the fine operator's effective wall
 * on every level.  No counterpart in the reference (its hierarchy is the
 * no-flux one, solvers/multigrid.py:905-1062); it changes the V-cycle
 * preconditioner only.  Refused (-1) unless every level runs on the tiled
 * kernels; the level operators of such a plan are solver-internal. */
int fvx_mg_plan_set_wall(void* plan, int on);
/* Chebyshev weights for the V-cycle's damped-Jacobi sweeps (both operators,
 * both precisions): sweep k of an m-sweep leg takes 1 / root_k of the
 * degree-m Chebyshev polynomial on [lo, 2] (the part of D^-1 A's spectrum
 * the smoother damps); lo <= 0 restores the uniform omega.  No reference
 * counterpart (multigrid.py:408-461 uses one weight); changes the V-cycle
 * preconditioner only (batched solves take each member's weights).
 * Refused (-1) unless the plan has 1-3 sweeps per leg, and the bottom takes
 * 1-3 sweeps on the tiled kernels or 1-8 in the single-workgroup coarse
 * kernel; the plan's level smoothing entry point is then solver-internal. */
int fvx_mg_plan_set_chebyshev(void* plan, double lo);
/* Level count, level shapes (2 ints per level) and the first level solved by
 * the single-workgroup coarse kernel. */
int fvx_mg_plan_info(void* plan, int* n_levels, int* shapes, int* coarse_level);
/* _apply_operator at one level (multigrid.py:327-405) */
int fvx_mg_apply_operator_f32(void* plan, int level, const float* u, float* out, void* stream);
int fvx_mg_apply_operator_f64(void* plan, int level, const double* u, double* out, void* stream);
/* _weighted_jacobi (multigrid.py:408-461) with the plan's omega; u NULL = 0 */
int fvx_mg_jacobi_f32(void* plan, int level, const float* u, const float* rhs, float* out,
                      int n_iters, void* stream);
int fvx_mg_jacobi_f64(void* plan, int level, const double* u, const double* rhs, double* out,
                      int n_iters, void* stream);
/* _restrict (multigrid.py:464-507) from `level` to level+1 */
int fvx_mg_restrict_f32(void* plan, int level, const float* v, float* out, void* stream);
int fvx_mg_restrict_f64(void* plan, int level, const double* v, double* out, void* stream);
/* _prolongate (multigrid.py:510-603) from level+1 onto `level` */
int fvx_mg_prolongate_f32(void* plan, int level, const float* vc, float* out, void* stream);
int fvx_mg_prolongate_f64(void* plan, int level, const double* vc, double* out, void* stream);
/* MultigridSolver.v_cycle at level 0 (multigrid.py:808-883); u NULL = 0.
 * make_multigrid_preconditioner (preconditioners.py:207-244) is V(0, r). */
int fvx_mg_v_cycle_f32(void* plan, const float* u, const float* rhs, float* out, void* stream);
int fvx_mg_v_cycle_f64(void* plan, const double* u, const double* rhs, double* out,
                       void* stream);
/* MultigridSolver.v_cycle(u, rhs, level_idx) (multigrid.py:808-883) from a
 * coarser level: u (NULL = 0), rhs and out have that level's shape.          */
int fvx_mg_v_cycle_level_f32(void* plan, int level, const float* u, const float* rhs,
                             float* out, void* stream);
int fvx_mg_v_cycle_level_f64(void* plan, int level, const double* u, const double* rhs,
                             double* out, void* stream);

typedef struct fvx_cg_info {
  int iterations;       /* inner PCG iterations (last refinement sweep for mixed) */
  int outer;            /* refinement sweeps (mixed precision), else 0 */
  int converged;        /* 1 if ||b - A x|| <= max(rtol ||b||, atol) */
  int pad;
  double residual_norm; /* ||b - A x||_2 (FP64 true residual for mixed solves) */
  double rhs_norm;      /* ||b||_2 */
} fvx_cg_info;

/* solve_cg(A = level-0 operator, preconditioner = V-cycle)
 * (solvers/iterative.py:48-136): flexible (Polak-Ribiere) PCG with device-side
 * reductions and convergence flags; rhs / x are pitched (row strides ld_*),
 * x is the warm start when use_x0 != 0.  Synchronises once per solve.      */
int fvx_mg_pcg_solve_f32(void* plan, const float* rhs, int64_t ld_rhs, float* x, int64_t ld_x,
                         double rtol, double atol, int max_iters, int use_x0,
                         fvx_cg_info* info, void* stream);
int fvx_mg_pcg_solve_f64(void* plan, const double* rhs, int64_t ld_rhs, double* x,
                         int64_t ld_x, double rtol, double atol, int max_iters, int use_x0,
                         fvx_cg_info* info, void* stream);
/* Same solve with FP32 multigrid-PCG inner solves and FP64 iterative
 * refinement (no reference counterpart: the precision scheme of this build).
 * Needs a plan created with precision 3264.                                */
int fvx_mg_mixed_solve(void* plan, const double* rhs, int64_t ld_rhs, double* x, int64_t ld_x,
                       double rtol, double atol, int max_outer, int max_inner, double inner_rtol,
                       fvx_cg_info* info, void* stream);
/* The same mixed solve with the inner preconditioner chosen by ``precond``:
 * 0 multigrid V-cycle (make_multigrid_preconditioner,
 * solvers/preconditioners.py:207-244), 1 learned U-Net ``net`` (from
 * fvx_unet_create; no reference counterpart), 2 diagonal (Jacobi).
 * out_scale <= 0 selects 1 / diag(A).  The learned preconditioner is a
 * two-grid cycle whose coarse solve is the network on multigrid level 1:
 * z0 = S^sweeps(0; r), res = r - A z0, c = Prolong(net(Restrict(res))),
 * z1 = z0 + alpha c with alpha = <res, c> / <c, A c>, z = S^sweeps(z1; r),
 * S the damped-Jacobi smoother of the plan; the network is applied as
 * out_scale rms(rc) net(rc / rms(rc)) and its plan must cover level 1.
 * info->pad counts inner solves that hit max_inner and were replaced by
 * the V-cycle.                                                            */
int fvx_mg_precond_solve(void* plan, int precond, void* net, double out_scale, int sweeps,
                         const double* rhs, int64_t ld_rhs, double* x, int64_t ld_x, double rtol,
                         double atol, int max_outer, int max_inner, double inner_rtol,
                         fvx_cg_info* info, void* stream);

/* Profiling hook (no reference counterpart; for benchmarks): reps
 * back-to-back launches of one fine-level PCG kernel on the plan's own
 * vectors, never converging.  which 0 = matvec a = A z with the PCG dots
 * (preconditioners other than the V-cycle), 1 = update (p = z + beta p,
 * q = a + beta q, x += alpha p, r -= alpha q), 2 = the level-0 up kernel of
 * a preconditioning V-cycle (prolongation of level 1, correction, the
 * post-sweeps, a = A z, PCG dots), 3 = the level-0 down kernel with the
 * previous PCG step applied first (all-wet FP32 Dirichlet plans); bits
 * 32/64 picks the hierarchy; reps = 0 prepares and must come first.        */
int fvx_mg_kernel_bench(void* plan, int which, int bits, int reps, void* stream);

/* ---- Work-list batches of independent solves (no reference counterpart:
 * batches the reference's per-layer vmap solves, solvers/elliptic.py:532-541,
 * and ensembles of models).  Members are mixed-precision plans (precision
 * 3264) with one operator kind and sweep count, each finest level tiled
 * (> 64 x 64 cells); every operator of the solve is ONE launch over all
 * members, each member guarded by its own convergence flags.               */
int fvx_mg_batch_create(void* const* plans, int n, void** batch_out);
int fvx_mg_batch_destroy(void* batch);
/* rhs[m], x[m]: member m's FP64 pitched right-hand side and warm start /
 * solution; n_outer refinement steps of n_inner[o] inner PCG iterations on
 * scalar slot ``slot``; enqueue only (the first call with a given array set
 * uploads its work-lists synchronously: make it before a graph capture). */
int fvx_mg_batch_mixed_enqueue(void* batch, const double* const* rhs, const int64_t* ld_rhs,
                               double* const* x, const int64_t* ld_x, double rtol, double atol,
                               int n_outer, const int* n_inner, double inner_rtol, int slot,
                               void* stream);
/* Per-member outcome of the last solve on ``slot`` (synchronises); counts
 * (nullable, 4 per member): inner iterations of each refinement step.      */
int fvx_mg_batch_status(void* batch, int slot, int n_outer_budget, fvx_cg_info* info,
                        int* counts, void* stream);

/* ---- Spectral rectangular Helmholtz solvers: solve_helmholtz_{dst,dct,fft}
 * (solvers/spectral.py:14-128, re-exported from spectraldiffx v0.0.10;
 * used by scripts/qg_1p5_layer.py:443).  bc 0 = DST-I (Dirichlet just
 * outside the array), 1 = DCT-II (no-flux cell-centred Neumann), 2 = FFT
 * (periodic); each is the exact inverse of the matching 5-point operator.
 * The plan owns hipFFT plans and work buffers for one (ny, nx, bc, dtype). */
int fvx_spectral_plan_create(int ny, int nx, int bc, int precision, void** plan_out);
int fvx_spectral_plan_destroy(void* plan);
int fvx_spectral_helmholtz_f32(void* plan, const float* rhs, int64_t ld_rhs, float* out,
                               int64_t ld_out, double dx, double dy, double lambda_,
                               void* stream);
int fvx_spectral_helmholtz_f64(void* plan, const double* rhs, int64_t ld_rhs, double* out,
                               int64_t ld_out, double dx, double dy, double lambda_,
                               void* stream);

/* ---- QG 1.5-layer step (composition; nearest reference
 * scripts/qg_1p5_layer.py:420-497; operators as cited above) -------------- */
typedef struct fvx_qg_info {
  int solves, iterations, max_iterations, outer, converged;
  int retries; /* graph steps redone with a wider iteration budget */
  double max_rel_residual;
} fvx_qg_info;
/* state_precision 32|64; solve_mode 64 (FP64 MG-PCG) | 3264 (mixed);
 * mask_interior: host float64 (ny, nx) or NULL (rectangular basin).       */
int fvx_qg_plan_create(int ny, int nx, double Lx, double Ly, double beta, double Ld, double nu4,
                       double F0, double dt, const double* mask_interior, int state_precision,
                       int solve_mode, int n_pre, int n_post, int n_coarse, double omega,
                       double rtol, double inner_rtol, int max_iters, int max_outer,
                       void** plan_out);
int fvx_qg_plan_destroy(void* plan);
/* Coarse-level wall correction (fvx_mg_plan_set_wall) of the inner FP32
 * V-cycles of a mixed-precision plan: *enabled = 1 when on, 0 when the plan's
 * hierarchy cannot take it or on = 0.  No reference counterpart; the FP64
 * certification is unchanged.  Invalidates captured graphs. */
int fvx_qg_set_wall(void* plan, int on, int* enabled);
/* Chebyshev smoothing (fvx_mg_plan_set_chebyshev) of the V-cycles of the
 * plan's solves (lo <= 0: uniform weight): *enabled = 1 when on.  A mixed
 * plan starts with it on where the hierarchy takes it (lo = 0.25 with a land
 * mask, 0.4 without); an FP64 plan starts with the uniform weight (the
 * reference's smoother).  No reference counterpart; the FP64 certification
 * is unchanged.  Invalidates captured graphs. */
int fvx_qg_set_chebyshev(void* plan, double lo, int* enabled);
/* nsteps rk3_ssp_step steps of q (full (ny+2, nx+2), state precision) in
 * place; psi (FP64 full array) is the warm start and the last solution.    */
int fvx_qg_step(void* plan, void* q, double* psi, int nsteps, fvx_qg_info* info, void* stream);
/* Learned U-Net preconditioner for the inner FP32 solves of a mixed plan
 * (NULL restores multigrid; no reference counterpart); out_scale <= 0 =
 * inverse Jacobi diagonal; the network is the coarse solve of the two-grid
 * cycle described at fvx_mg_precond_solve, with ``sweeps`` pre/post sweeps.
 * Multigrid-PCG takes over when a learned inner solve hits its cap;
 * fvx_qg_fallbacks counts those events. */
int fvx_qg_set_preconditioner(void* plan, void* net, double out_scale, int sweeps);
int fvx_qg_fallbacks(void* plan);
/* Budgets of the captured step graphs: out[0] FP64 PCG iterations, out[1]
 * inner FP32 iterations per refinement sweep, out[2] refinement sweeps,
 * out[3] recaptures so far (no reference counterpart).                    */
int fvx_qg_budgets(void* plan, int* out);
/* fvx_qg_step replaying one captured HIP graph per step (three solves with
 * fixed iteration budgets and device-side convergence guards, FP64 residual
 * certification, three stage kernels); a step whose solve misses rtol is
 * redone with a wider budget.  Budgets are calibrated by the first call. */
int fvx_qg_step_graph(void* plan, void* q, double* psi, int nsteps, fvx_qg_info* info,
                      void* stream);
/* A fixed work-list of QG stage evaluations over many FP64 fields (ensemble
 * members, layers): one launch evaluates every entry; tend32 computes the
 * tendencies in FP32 (no reference counterpart).                          */
typedef struct fvx_qg_stage_desc {
  const double* qs;
  const double* psi;
  const double* qn;
  double* out;
  const uint8_t* mask; /* full-array wet mask or NULL */
  int Ny, Nx;
  double dx, dy, beta, Ly, F0, nu4, a, b, cdt;
} fvx_qg_stage_desc;
int fvx_qg_stage_batch_create(int n, const fvx_qg_stage_desc* descs, int tend32, void** out);
int fvx_qg_stage_batch_launch(void* handle, void* stream);
int fvx_qg_stage_batch_destroy(void* handle);
/* ---- Geostrophic face velocities of a T-point streamfunction ---------------
 * geostrophic_velocity_from_streamfunction (scripts/qg_1p5_layer.py:115-145):
 * psi averaged to X points (Interpolation2D.T_to_X, interpolation.py:158-168),
 * u = -diff_y_X_to_U, v = diff_x_X_to_V (difference.py:280-320), interior
 * only with a zero ring; with masks (all three or none) the masked operators'
 * xy_corner_strict / u / v multiply each result.  psi is FP64; u, v the
 * suffix's type (a tracer's).                                               */
int fvx_geostrophic_faces_f32(const double* psi, const uint8_t* mask_xy, const uint8_t* mask_u,
                              const uint8_t* mask_v, float* u, float* v, int ny, int nx,
                              double dx, double dy, void* stream);
int fvx_geostrophic_faces_f64(const double* psi, const uint8_t* mask_xy, const uint8_t* mask_u,
                              const uint8_t* mask_v, double* u, double* v, int ny, int nx,
                              double dx, double dy, void* stream);
/* One fused tendency + stage combination out = a q_n + b q_s + cdt rhs(q_s, psi). */
int fvx_qg_stage_f32(const float* qs, const double* psi, const float* qn, float* out,
                     const uint8_t* mask, int Ny, int Nx, double dx, double dy, double beta,
                     double Ly, double F0, double nu4, double a, double b, double cdt,
                     void* stream);
int fvx_qg_stage_f64(const double* qs, const double* psi, const double* qn, double* out,
                     const uint8_t* mask, int Ny, int Nx, double dx, double dy, double beta,
                     double Ly, double F0, double nu4, double a, double b, double cdt,
                     void* stream);
/* FP64 state with the tendency computed in FP32 (the mixed-precision step). */
int fvx_qg_stage_f64t32(const double* qs, const double* psi, const double* qn, double* out,
                        const uint8_t* mask, int Ny, int Nx, double dx, double dy, double beta,
                        double Ly, double F0, double nu4, double a, double b, double cdt,
                        void* stream);
/* The FP64-state stage with the certification residual of the Helmholtz solve
 * (lap - lam) psi = q_s (interior, masked Dirichlet operator) fused in: block
 * partials of ||q_s m||^2 into partials[0 .. *nblk-1] and of ||r||^2 into
 * partials[*nblk .. 2 *nblk - 1] (device array of partials_cap doubles). */
int fvx_qg_stage_cert_f64(const double* qs, const double* psi, const double* qn, double* out,
                          const uint8_t* mask, int Ny, int Nx, double dx, double dy, double beta,
                          double Ly, double F0, double nu4, double a, double b, double cdt,
                          double lam, double* partials, int64_t partials_cap, int* nblk,
                          void* stream);
int fvx_qg_stage_cert_f64t32(const double* qs, const double* psi, const double* qn, double* out,
                             const uint8_t* mask, int Ny, int Nx, double dx, double dy,
                             double beta, double Ly, double F0, double nu4, double a, double b,
                             double cdt, double lam, double* partials, int64_t partials_cap,
                             int* nblk, void* stream);

/* ---- Learned U-Net preconditioner (no reference counterpart) -------------
 * Bias-free U-Net: nlev levels of two 3x3 convs + ReLU (channels[l], each a
 * multiple of 16 <= 64), 2x2 average pooling, nearest upsampling, additive
 * skips, 1-channel input/output; BF16 weights/activations, FP32 accumulation
 * on matrix cores; FP32 master weights trained with Adam.  The CPU oracle is
 * oracle/unet.py.  H, W: largest field the plan runs (divisible by 2^(nlev-1)). */
int fvx_unet_create(int nlev, const int* channels, int H, int W, void** out);
int fvx_unet_destroy(void* net);
int64_t fvx_unet_num_weights(void* net);
/* direction 0: load fp32 weights from device memory (resets Adam); 1: read. */
int fvx_unet_weights(void* net, float* w, int direction, void* stream);
int fvx_unet_forward_raw(void* net, const float* x, float* y, int H, int W, void* stream);
int fvx_unet_backward_raw(void* net, const float* x, const float* gy, int H, int W,
                          void* stream);
int fvx_unet_grads(void* net, float* g, int zero_after, void* stream);
/* Adam step on a batch of B (x, e) pairs, loss = mean (net(x) - e)^2. */
int fvx_unet_train_step(void* net, const float* xb, const float* eb, int B, int H, int W,
                        double lr, double* loss_out, void* stream);
/* z = out_scale * rms(r) * net(r / rms(r)) * mask  (the PCG preconditioner). */
int fvx_unet_apply_f32(void* net, const float* r, int64_t ld_r, float* z, int64_t ld_z, int H,
                       int W, double out_scale, const uint8_t* mask, void* stream);
/* One 3x3 conv on matrix cores (bf16 channels-last in/out, fp32 weights). */
int fvx_conv3x3_bf16(const uint16_t* x, int Cin, const float* w, int Cout, uint16_t* out, int H,
                     int W, int ups, int relu, uint16_t* scratch, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FVX_H */
