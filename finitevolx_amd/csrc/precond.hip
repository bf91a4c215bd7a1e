// Preconditioner choice for the mixed-precision PCG Helmholtz solve: the
// multigrid V-cycle (the reference's make_multigrid_preconditioner,
// solvers/preconditioners.py:207-244), the learned U-Net (as the coarse solve
// of a two-grid cycle), or the diagonal (Jacobi, z = r / diag(A) on wet
// cells).  The diagonal option is the
// baseline a learned preconditioner has to beat; the V-cycle is the one it
// has to match.

#include "fvx_common.hpp"
#include "mg.hpp"

namespace fvx {
template <typename T>
int unet_apply(void* net, const T* r, int64_t ld_r, T* z, int64_t ld_z, int H, int W,
               double out_scale, const uint8_t* mask, hipStream_t s, const int* guard);

__global__ __launch_bounds__(256) void diag_precond_kernel(int ny, int nx, float scale,
                                                           const uint8_t* __restrict__ mask,
                                                           const float* __restrict__ r,
                                                           float* __restrict__ z,
                                                           const int* __restrict__ guard) {
  if (guard && *guard) return;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int64_t k = (int64_t)j * nx + i;
  z[k] = (mask && !mask[k]) ? 0.f : scale * r[k];
}

struct PrecondCtx {
  Plan* plan;
  void* net;
  double scale;
  int sweeps;
};

struct NetCtx {
  void* net;
  double scale;
};

static int coarse_net(void* ctx, const float* rc, float* zc, int nyc, int nxc,
                      const uint8_t* maskc, const int* guard, hipStream_t s) {
  NetCtx* c = (NetCtx*)ctx;
  return unet_apply<float>(c->net, rc, nxc, zc, nxc, nyc, nxc, c->scale, maskc, s, guard);
}

// Learned preconditioner: two-grid cycle whose coarse solve is the network
// (damped-Jacobi pre/post sweeps on the fine level, network on level 1,
// energy-minimising scaling of the coarse correction).
int learned_precond_apply(Plan& P, void* net, double scale, int sweeps, const float* r, float* z,
                          const int* guard, hipStream_t s) {
  NetCtx ctx{net, scale};
  return learned_two_grid(P, &coarse_net, &ctx, sweeps, r, z, guard, s);
}

static int learned_fn(void* ctx, const float* r, float* z, int ny, int nx, const int* guard,
                      hipStream_t s) {
  (void)ny;
  (void)nx;
  PrecondCtx* c = (PrecondCtx*)ctx;
  return learned_precond_apply(*c->plan, c->net, c->scale, c->sweeps, r, z, guard, s);
}

static int diag_fn(void* ctx, const float* r, float* z, int ny, int nx, const int* guard,
                   hipStream_t s) {
  PrecondCtx* c = (PrecondCtx*)ctx;
  hipLaunchKernelGGL(diag_precond_kernel, dim3(cdiv(nx, 64), cdiv(ny, 4)), dim3(64, 4), 0, s, ny,
                     nx, (float)c->scale, plan_mask0(*c->plan), r, z, guard);
  FVX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace fvx

extern "C" {

// Mixed-precision PCG (FP32 inner flexible PCG, FP64 refinement) with the
// inner preconditioner chosen by ``precond``: 0 multigrid V-cycle, 1 learned
// two-grid cycle (``sweeps`` damped-Jacobi pre/post sweeps; coarse solve
// out_scale rms(rc) net(rc / rms(rc)) on level 1, scaled to minimise the
// energy norm), 2 diagonal (out_scale <= 0 selects 1 / diag(A)).  info->pad returns the
// number of learned/diagonal inner solves that hit max_inner and were
// replaced by the V-cycle.
int fvx_mg_precond_solve(void* plan, int precond, void* net, double out_scale, int sweeps,
                         const double* rhs, int64_t ld_rhs, double* x, int64_t ld_x, double rtol,
                         double atol, int max_outer, int max_inner, double inner_rtol,
                         void* info, void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(plan && rhs && x, "null argument");
  FVX_CHECK_ARG(precond >= 0 && precond <= 2, "precond must be 0 (multigrid), 1 (learned) or 2 "
                                              "(diagonal)");
  FVX_CHECK_ARG(precond != 1 || net, "learned preconditioner needs a network");
  int nyc = 0, nxc = 0;
  FVX_CHECK_ARG(precond != 1 || learned_coarse_dims(*(Plan*)plan, &nyc, &nxc) == 0,
                "the learned preconditioner needs a mixed plan with at least two levels");
  Plan* P = (Plan*)plan;
  FVX_CHECK_ARG(plan_has_mixed(*P), "mixed solve needs a plan built with precision 3264");
  FVX_CHECK_ARG(max_outer >= 1 && max_inner >= 1, "iteration caps must be >= 1");
  FVX_CHECK_ARG(ld_rhs >= plan_nx(*P) && ld_x >= plan_nx(*P), "row pitch smaller than nx");
  FVX_CHECK_ARG(sweeps >= 0 && sweeps <= 8, "sweeps must be in [0, 8]");
  PrecondCtx ctx{P, net, out_scale > 0 ? out_scale : plan_inv_diag0(*P), sweeps};
  PrecondFn pf = precond == 1 ? &learned_fn : precond == 2 ? &diag_fn : nullptr;
  return mixed_solve<double>(*P, rhs, ld_rhs, x, ld_x, rtol, atol, max_outer, max_inner,
                             inner_rtol, (CgInfoC*)info, 0, (hipStream_t)stream, pf,
                             pf ? &ctx : nullptr);
}

}  // extern "C"
