// Internal interface between the multigrid/PCG module and its in-library
// callers (the QG step driver).  Not part of the C ABI.
#pragma once

#include "fvx_common.hpp"

namespace fvx {

// Device-side scalars of one PCG solve (one "slot" holds an inner and an
// outer block).
struct PcgScalars {
  double rz;      // rho_k = r . z
  double pq;      // p . A p
  double rr;      // ||r||^2 (recursive residual)
  double bb;      // ||b||^2
  double tol2;    // stop when rr <= tol2
  double alpha, beta;
  double zq;      // z_{k+1} . q_k  (flexible beta)
  double out_rr;  // outer (FP64) residual ||b - A x||^2 for refinement
  double out_tol2;
  int iter;       // inner iterations done
  int done;       // 1 once rr <= tol2 (all kernels become no-ops)
  int first;      // 1 before the first direction is formed
  int out_done;   // outer refinement converged
  int outer;      // outer iterations done
  int total;      // inner iterations summed over refinement steps
  int max_inner;  // largest inner iteration count of one refinement step
  int it_outer[4];  // inner iterations of refinement steps 0..3 (last solve)
  int pad;
};

struct CgInfoC {
  int iterations, outer, converged, pad;
  double residual_norm, rhs_norm;
};

struct Plan;

// A preconditioner replacing the V-cycle inside flexible PCG: z = M(r) on the
// finest level (FP32 arrays of nx columns); kernels should exit early when
// *guard is nonzero (the solve has converged).
using PrecondFn = int (*)(void* ctx, const float* r, float* z, int ny, int nx, const int* guard,
                          hipStream_t s);

template <typename T>
int pcg_enqueue(Plan& P, const T* rhs, int64_t ld_rhs, T* x, int64_t ld_x, double rtol,
                double atol, bool use_x0, int iters, int slot, hipStream_t s);
template <typename T>
int certify_enqueue(Plan& P, const T* rhs, int64_t ld_rhs, const T* x, int64_t ld_x, double rtol,
                    double atol, int slot, hipStream_t s);
template <typename T>
// rtol: the recursive residual's target; cert_rtol (< 0: rtol) the tolerance
// the FP64 true residual is certified against.
int pcg_solve(Plan& P, const T* rhs, int64_t ld_rhs, T* x, int64_t ld_x, double rtol,
              double atol, int max_iters, bool use_x0, CgInfoC* info, int slot, hipStream_t s,
              double cert_rtol = -1.0);
// defer_final: leave the closing FP64 residual to the caller, which writes
// its ||b||^2 and ||b - A x||^2 block partials into plan_partials() and calls
// certify_finish (the QG step does this inside its stage kernel).
template <typename TB>
int mixed_enqueue(Plan& P, const TB* rhs, int64_t ld_rhs, double* x, int64_t ld_x, double rtol,
                  double atol, int n_outer, const int* n_inner, double inner_rtol, int slot,
                  hipStream_t s, PrecondFn pf = nullptr, void* ctx = nullptr,
                  bool defer_final = false);
// The plan's reduction scratch and its capacity in doubles.
double* plan_partials(Plan& P, size_t* capacity);
// Finishing kernel of a certification residual whose block partials
// (value-major: ||b||^2 of block k at [k], ||r||^2 at [nblk + k]) a caller has
// written: the mixed solve's closing residual (fp64_true = false) or the
// FP64 PCG's true residual (fp64_true = true) of slot ``slot``.
void certify_finish(Plan& P, int nblk, double rtol, double atol, int slot, bool fp64_true,
                    hipStream_t s);
// Synchronous mixed solve.  With ``pf`` the inner FP32 solves use it; an
// inner solve that hits ``max_inner`` without reaching its tolerance makes the
// remaining sweeps fall back to the multigrid V-cycle (counted in
// info->pad).
template <typename TB>
// info->pad returns the number of learned inner solves replaced by the
// V-cycle; plan_max_inner then gives the largest iteration count of one inner
// solve.
int mixed_solve(Plan& P, const TB* rhs, int64_t ld_rhs, double* x, int64_t ld_x, double rtol,
                double atol, int max_outer, int max_inner, double inner_rtol, CgInfoC* info,
                int slot, hipStream_t s, PrecondFn pf = nullptr, void* ctx = nullptr);
const uint8_t* plan_mask0(const Plan& P);
double plan_inv_diag0(const Plan& P);

// Copy ``n`` slots of device scalars to the plan's pinned mirror and wait;
// returns the host pointer (2 blocks per slot).
const PcgScalars* fetch_slots(Plan& P, int first_slot, int n, hipStream_t s);
int plan_levels(const Plan& P);
// Coarse-level wall correction of an FP32 Dirichlet hierarchy (level l's
// dry or outside neighbours take the ghost value g_l u, g_l = (1 - 2^l) /
// (1 + 2^l)); refused unless every level runs on the wave-tile kernels.
int plan_set_wall(Plan& P, bool on);
// Chebyshev weights for the wave-tile V-cycle's sweeps (smoothing interval
// [lo, 2] of D^-1 A; lo <= 0: the uniform omega); refused unless every level
// runs on the tiled kernels with 1-3 sweeps per leg and at the bottom.
int plan_set_chebyshev(Plan& P, double lo);
bool plan_has_mixed(const Plan& P);
// Learned two-grid preconditioner on the FP32 hierarchy, z = M(r):
//   z0 = S^sweeps(0; r), res = r - A z0          (damped-Jacobi smoother S)
//   c  = Prolong(net(Restrict(res)))              (network on level 1)
//   z1 = z0 + alpha c, alpha = <res, c> / <c, A c> (energy-minimising)
//   z  = S^sweeps(z1; r)
// ``fn`` applies the network to a level-1 field (contiguous, nyc x nxc).
typedef int (*CoarseNetFn)(void* ctx, const float* rc, float* zc, int nyc, int nxc,
                           const uint8_t* maskc, const int* guard, hipStream_t s);
int learned_two_grid(Plan& P, CoarseNetFn fn, void* ctx, int sweeps, const float* r, float* z,
                     const int* guard, hipStream_t s);
// Level-1 dimensions of the FP32 hierarchy (the network's grid).
int learned_coarse_dims(const Plan& P, int* nyc, int* nxc);
// Lazily allocated fine-level FP32 work arrays k = 0..3.
float* fine_work(Plan& P, int k);
int plan_nx(const Plan& P);
int plan_ny(const Plan& P);
// Largest inner iteration count of one refinement step of the last
// synchronous mixed solve on ``slot``; plan_inner_iters fills per_outer[o]
// (o < n) with refinement step o's inner iterations and returns the number
// of steps.
int plan_max_inner(const Plan& P, int slot);
// The synchronous mixed solve's opening schedule (refinement steps x inner
// iterations, each capped by the call's limits); it adapts to the last solve.
void plan_schedule(const Plan& P, int* outer, int* inner);
void plan_set_schedule(Plan& P, int outer, int inner);
int plan_inner_iters(const Plan& P, int slot, int* per_outer, int n);
// Learned preconditioner application (precond.hip); sweeps as in
// fvx_mg_precond_solve.
int learned_precond_apply(Plan& P, void* net, double scale, int sweeps, const float* r, float* z,
                          const int* guard, hipStream_t s);
int plan_kind(const Plan& P);

}  // namespace fvx
