// Internal interface between the multigrid/PCG module and its in-library
// callers (the QG step driver).  Not part of the C ABI.
#pragma once

#include "fvx_common.hpp"

namespace fvx {

// Counters per reduction site: one top level + one per block group.
constexpr int kRedGroups = 64;

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
  int pad[2];
  unsigned int ticket[8][kRedGroups + 1];  // last-block counters per reduction site
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
int pcg_solve(Plan& P, const T* rhs, int64_t ld_rhs, T* x, int64_t ld_x, double rtol,
              double atol, int max_iters, bool use_x0, CgInfoC* info, int slot, hipStream_t s);
template <typename TB>
int mixed_enqueue(Plan& P, const TB* rhs, int64_t ld_rhs, double* x, int64_t ld_x, double rtol,
                  double atol, int n_outer, int n_inner, double inner_rtol, int slot,
                  hipStream_t s, PrecondFn pf = nullptr, void* ctx = nullptr);
// Synchronous mixed solve.  With ``pf`` the inner FP32 solves use it; an
// inner solve that hits ``max_inner`` without reaching its tolerance makes the
// remaining sweeps fall back to the multigrid V-cycle (counted in
// info->pad).
template <typename TB>
int mixed_solve(Plan& P, const TB* rhs, int64_t ld_rhs, double* x, int64_t ld_x, double rtol,
                double atol, int max_outer, int max_inner, double inner_rtol, CgInfoC* info,
                int slot, hipStream_t s, PrecondFn pf = nullptr, void* ctx = nullptr);
const uint8_t* plan_mask0(const Plan& P);
double plan_inv_diag0(const Plan& P);

// Copy ``n`` slots of device scalars to the plan's pinned mirror and wait;
// returns the host pointer (2 blocks per slot).
const PcgScalars* fetch_slots(Plan& P, int first_slot, int n, hipStream_t s);
int plan_levels(const Plan& P);
bool plan_has_mixed(const Plan& P);
// Fine-level FP32 helpers (level 0 of the FP32 hierarchy, contiguous fields):
// lazily allocated work arrays k = 0..2; ``sweeps`` damped-Jacobi sweeps on
// A z = r from zero with res = (r - A z) mask; and sweeps from z_in (+ add)
// into z_out.  z_in is overwritten when ``add`` is given.
float* fine_work(Plan& P, int k);
int fine_presmooth(Plan& P, const float* r, float* z, float* res, int sweeps, hipStream_t s);
int fine_postsmooth(Plan& P, float* z_in, const float* add, const float* r, float* z_out,
                    int sweeps, hipStream_t s);
int plan_nx(const Plan& P);
int plan_ny(const Plan& P);
// Learned preconditioner application (precond.hip); sweeps as in
// fvx_mg_precond_solve.
int learned_precond_apply(Plan& P, void* net, double scale, int sweeps, const float* r, float* z,
                          const int* guard, hipStream_t s);
int plan_kind(const Plan& P);

}  // namespace fvx
