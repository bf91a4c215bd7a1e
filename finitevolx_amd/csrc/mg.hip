// Geometric multigrid for the Helmholtz operator and multigrid-preconditioned
// flexible CG, FP64, FP32, and FP32 with FP64 iterative refinement.
//
// Reference semantics (finitevolx 0.0.40):
//  * hierarchy build: build_multigrid_solver (solvers/multigrid.py:905-1062)
//  * operator:        _apply_operator (:327-405)        kind FVX_MG_NOFLUX
//                     masked_laplacian (iterative.py:144-191) on a
//                     zero-padded array                 kind FVX_MG_DIRICHLET
//  * smoother:        _weighted_jacobi (:408-461)
//  * transfers:       _restrict (:464-507), _prolongate (:510-603)
//  * V-cycle:         MultigridSolver.v_cycle (:808-883)
//  * preconditioner:  make_multigrid_preconditioner (preconditioners.py:207-244)
//
// GPU design (gfx950): per fine level one fused DOWN kernel (nu1 Jacobi sweeps
// temporally blocked in an LDS tile + residual + 2x2 restriction) and one fused
// UP kernel (prolongation + correction + nu2 sweeps); the coarse levels
// (<= kCoarseCells cells) run inside a single workgroup; PCG vector updates
// and dot products are fused kernels whose last workgroup finishes each
// reduction on the device (no host round trip); every kernel of an iteration
// exits immediately once the device-side convergence flag is set, so the host
// can enqueue a predicted number of iterations and synchronise once.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <list>
#include <vector>

#include "fvx_common.hpp"
#include "mg.hpp"

namespace fvx {

enum MgKind : int { kNoflux = 0, kDirichlet = 1 };

template <typename T>
struct LevelDev {
  int ny, nx;
  T idx2, idy2, lam;
  T inv_d;                 // Dirichlet scalar 1/diag
  const uint8_t* mask;     // nullptr = all wet
  const T* cx;             // noflux face coefficients (nullptr for Dirichlet)
  const T* cy;
  const T* inv;            // noflux inverse diagonal
  T* u;                    // iterate (coarse levels: error)
  T* b;                    // right-hand side (coarse levels: restricted residual)
  T* t;                    // scratch (ping-pong for single-workgroup sweeps)
  T* t2;                   // scratch for tiled sweep chains
  T* t3;
  // Dirichlet coarse levels with the wall correction: each dry or outside
  // neighbour adds wall_g / h^2 to the diagonal (a ghost value wall_g u_c
  // instead of 0); 0 = off.
  T wall_g;
  // The mask packed by columns (with mask): word (blk, i) = the mask of rows
  // 32 blk .. 32 blk + 31 of column i as bits (rows beyond the level 0),
  // blocks 0 .. ny / 32 + 1, so a 64-bit window from any row block is valid.
  const uint32_t* mcol;
};

template <typename T>
__device__ __forceinline__ T mval(const uint8_t* m, int64_t k) {
  return m ? (m[k] ? T(1) : T(0)) : T(1);
}

// ---------------------------------------------------------------------------
// Reductions: every block reduces its values in registers and LDS and writes
// one partial per value with plain stores; a one-workgroup finishing kernel
// launched right after sums the partials in block order (deterministic) and
// updates the solver scalars on the device.  The kernel boundary orders the
// partial stores before their reads, so no per-block device-scope fence is
// needed (on this GPU such a fence writes back the XCD's L2: with one per
// block a 16k-block reduction cost ~0.4 ms).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lin_tid() {
  return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}

// Sums of NV values over the block into partials: block ``bid`` of ``nblk``
// stores value k at partials[k * nblk + bid].
template <int NV, int kThreads, typename A>
__device__ __forceinline__ void block_partials(const A (&v)[NV], double* partials, int64_t nblk,
                                               int64_t bid) {
  static_assert(NV <= 64, "one summing thread per value");
  __shared__ double red[kThreads / 64][NV];
  const int tid = lin_tid();
  const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const A s = wave_sum_last(v[k]);
    if (lane == 63) red[wid][k] = (double)s;
  }
  __syncthreads();
  if (tid < NV) {
    // Value-major: value k of block b at partials[k * nblocks + b], so the
    // finishing kernel's reads are coalesced.  One thread per value sums the
    // waves in order.
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) s += red[w][tid];
    partials[tid * nblk + bid] = s;
  }
}
template <int NV, int kThreads, typename A>
__device__ __forceinline__ void block_partials(const A (&v)[NV], double* partials) {
  block_partials<NV, kThreads>(v, partials, (int64_t)gridDim.x * gridDim.y,
                               (int64_t)blockIdx.y * gridDim.x + blockIdx.x);
}

// Reduction sites (which scalars the finishing kernel updates).
enum RedSite : int {
  kRedCG = 0,      // the nine PCG dots -> beta, p.Ap, alpha, predicted ||r||^2
  kRedInit = 1,    // ||b||^2, ||r0||^2 -> tolerance, flags
  kRedRefine = 4,  // FP64 refinement residual
  kRedRefineReset = 5,  // the same, then the inner solve's reset (sc - 1)
  kRedTrue = 6,    // FP64 certification residual
};

// Reset of an inner FP32 solve on an already-formed refinement residual:
// tolerance from ||r0||, flags cleared; single thread.  The refinement step
// that has just finished keeps its iteration count.  Reduce the residual by
// inner_rtol, but no further than half the outer target: later refinement
// steps then stop as soon as the FP64 residual can certify (FP32 rounding of
// the correction is ~eps32 kappa ||r||, far below that target once ||r|| is
// small).
__device__ __forceinline__ void inner_reset(PcgScalars* inner, const PcgScalars* outer,
                                            double inner_rtol) {
  if (outer->outer >= 1 && outer->outer <= 4) inner->it_outer[outer->outer - 1] = inner->iter;
  if (outer->out_done) {
    inner->done = 1;
    return;
  }
  inner->max_inner = max(inner->max_inner, inner->iter);
  inner->rr = outer->out_rr;
  inner->tol2 = fmax(inner_rtol * inner_rtol * outer->out_rr, 0.25 * outer->out_tol2);
  inner->done = 0;
  inner->first = 1;
  inner->iter = 0;
}
// Order of the PCG dots (z = M r, a = A z, q = q_{k-1} = A p_{k-1}).
enum CgDot : int { kZR, kZQ, kZA, kRR, kRA, kRQ, kAA, kAQ, kQQ, kNumCgDots };
constexpr int kFinishThreads = 1024;

template <int NV, int SITE>
__device__ __forceinline__ void reduce_finish_body(const double* __restrict__ partials,
                                                   int nblk, PcgScalars* sc, double rtol2,
                                                   double atol2, double inner_rtol) {
  // kRedRefineReset: the inner solve's scalars sit just before the outer ones.
  constexpr bool kReset = SITE == kRedRefineReset;
  // The same guards as the producing kernels, read together with the
  // partials so that the two memory round trips overlap.
  int skip = 0;
  if constexpr (SITE == kRedRefine || kReset) skip = sc->out_done;
  else if constexpr (SITE != kRedInit && SITE != kRedTrue) skip = sc->done;
  // The PCG step's previous-iteration scalars, in the same round trip.
  int first = 1;
  double pq_prev = 0.0, tol2 = 0.0;
  if constexpr (SITE == kRedCG) {
    first = sc->first;
    pq_prev = sc->pq;
    tol2 = sc->tol2;
  }
  __shared__ double red[kFinishThreads / 64][NV];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
#pragma unroll 4
  for (int b = tid; b < nblk; b += kFinishThreads) {
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += partials[(int64_t)k * nblk + b];
  }
  if (skip) {
    // done == 2: the last update has been applied; stop it being applied again.
    if constexpr (SITE == kRedCG)
      if (tid == 0 && skip == 2) sc->done = 1;
    if constexpr (kReset)
      if (tid == 0) inner_reset(sc - 1, sc, inner_rtol);
    return;
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = wave_sum_last(acc[k]);
    if (lane == 63) red[wid][k] = s;
  }
  __syncthreads();
  // One thread per value sums the waves in order (the LDS reads of all values
  // overlap), then thread 0 applies the site's update.
  __shared__ double tot[NV];
  if (tid < NV) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kFinishThreads / 64; ++w) s += red[w][tid];
    tot[tid] = s;
  }
  __syncthreads();
  if (tid != 0) return;
  double out[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = tot[k];
  if constexpr (SITE == kRedCG) {
    // One flexible-PCG step from the dots of z = M r_k and a = A z_k (the
    // preconditioned CG that solve_cg, iterative.py:48-141, takes from
    // lineax.CG, with the Polak-Ribiere beta of flexible PCG):
    //   beta  = -z.q_{k-1} / p_{k-1}.q_{k-1}   (Polak-Ribiere: z.(r_k - r_{k-1}) / rho_{k-1})
    //   p_k = z + beta p_{k-1},  q_k = A p_k = a + beta q_{k-1}
    //   p_k.q_k = z.a + 2 beta z.q_{k-1} + beta^2 p_{k-1}.q_{k-1}   (A symmetric)
    //   alpha = r_k.z / p_k.q_k
    //   ||r_{k+1}||^2 = ||r_k - alpha q_k||^2 expanded in the same dots.
    // The update kernel applies the step; when the predicted residual meets
    // the tolerance, done = 2 lets that last update run and stops the rest.
    double beta = first ? 0.0 : -out[kZQ] / pq_prev;
    double pq = out[kZA] + beta * (2.0 * out[kZQ] + beta * (first ? 0.0 : pq_prev));
    // p.Ap = z.Az (1 - cos^2) with cos the A-angle between z and p_{k-1}:
    // when that nearly cancels, restart from p = z.
    if (beta != 0.0 && !(fabs(pq) > 1e-6 * fabs(out[kZA]))) {
      beta = 0.0;
      pq = out[kZA];
    }
    // (A and M may have opposite signs -- PCG is invariant to that.)  Zero or
    // non-finite: breakdown, no update; the caller certifies the residual.
    if (!(fabs(pq) > 0.0 && fabs(pq) < HUGE_VAL)) {
      sc->done = 1;
      return;
    }
    const double alpha = out[kZR] / pq;
    const double rq = out[kRA] + beta * out[kRQ];
    const double qq = out[kAA] + beta * (2.0 * out[kAQ] + beta * out[kQQ]);
    const double rr = out[kRR] - alpha * (2.0 * rq - alpha * qq);
    sc->beta = beta;
    sc->pq = pq;
    sc->alpha = alpha;
    sc->rz = out[kZR];
    sc->zq = out[kZQ];
    sc->rr = rr;
    sc->first = 0;
    sc->iter += 1;
    sc->total += 1;
    if (rr <= tol2) sc->done = 2;
  } else if constexpr (SITE == kRedInit) {
    sc->bb = out[0];
    sc->rr = out[1];
    sc->tol2 = fmax(rtol2 * out[0], atol2);
    sc->done = out[1] <= sc->tol2 ? 1 : 0;
    sc->first = 1;
    sc->iter = 0;
    sc->total = 0;
    sc->alpha = 0.0;
    sc->beta = 0.0;
  } else {  // kRedRefine, kRedRefineReset, kRedTrue
    sc->bb = out[0];
    sc->out_rr = out[1];
    sc->out_tol2 = fmax(rtol2 * out[0], atol2);
    sc->out_done = out[1] <= sc->out_tol2 ? 1 : 0;
    if constexpr (kReset) inner_reset(sc - 1, sc, inner_rtol);
  }
}
template <int NV, int SITE>
__global__ __launch_bounds__(kFinishThreads) void reduce_finish_kernel(
    const double* __restrict__ partials, int nblk, PcgScalars* sc, double rtol2, double atol2,
    double inner_rtol = 0.0) {
  reduce_finish_body<NV, SITE>(partials, nblk, sc, rtol2, atol2, inner_rtol);
}

// ---------------------------------------------------------------------------
// Operator at a cell of an LDS tile (u, cx, cy indexed by tile coordinates).
// ---------------------------------------------------------------------------
template <typename T, int KIND, int W>
__device__ __forceinline__ T tile_Au(const T* U, const T* CX, const T* CY, T m, int y, int x,
                                     T idx2, T idy2, T lam) {
  const int k = y * W + x;
  const T u = U[k];
  T lap;
  if constexpr (KIND == kDirichlet) {
    lap = (U[k + 1] + U[k - 1] - T(2) * u) * idx2 + (U[k + W] + U[k - W] - T(2) * u) * idy2;
  } else {
    lap = (CX[k] * (U[k + 1] - u) - CX[k - 1] * (u - U[k - 1])) * idx2 +
          (CY[k] * (U[k + W] - u) - CY[k - W] * (u - U[k - W])) * idy2;
  }
  return (lap - lam * u) * m;
}

// Operator at a cell of a global level array (zero outside the array).
template <typename T, int KIND>
__device__ __forceinline__ T glob_Au(const LevelDev<T>& L, const T* u, int j, int i) {
  const int64_t k = (int64_t)j * L.nx + i;
  const T m = mval<T>(L.mask, k);
  // The reference applies the stencil to u * mask (multigrid.py:383,
  // iterative.py:183); with face coefficients that vanish across land only
  // the Dirichlet operator needs the neighbours masked.
  auto U = [&](int64_t kk) {
    return KIND == kDirichlet ? u[kk] * mval<T>(L.mask, kk) : u[kk];
  };
  const T c = u[k] * m;
  const T e = (i + 1 < L.nx) ? U(k + 1) : T(0);
  const T w = (i > 0) ? U(k - 1) : T(0);
  const T n = (j + 1 < L.ny) ? U(k + L.nx) : T(0);
  const T s = (j > 0) ? U(k - L.nx) : T(0);
  T lap;
  if constexpr (KIND == kDirichlet) {
    lap = (e + w - T(2) * c) * L.idx2 + (n + s - T(2) * c) * L.idy2;
  } else {
    const T cxe = L.cx[k], cxw = (i > 0) ? L.cx[k - 1] : T(0);
    const T cyn = L.cy[k], cys = (j > 0) ? L.cy[k - L.nx] : T(0);
    lap = (cxe * (e - c) - cxw * (c - w)) * L.idx2 + (cyn * (n - c) - cys * (c - s)) * L.idy2;
  }
  return (lap - L.lam * c) * m;
}

// Operator on a pitched array (row stride ld), zero outside the level.
template <typename T, int KIND>
__device__ __forceinline__ T glob_Au_pitched(const LevelDev<T>& L, const T* u, int64_t ld, int j,
                                             int i) {
  const int64_t k = (int64_t)j * L.nx + i;
  const int64_t ku = (int64_t)j * ld + i;
  const T m = mval<T>(L.mask, k);
  auto U = [&](int64_t kk, int64_t km) {
    return KIND == kDirichlet ? u[kk] * mval<T>(L.mask, km) : u[kk];
  };
  const T c = u[ku] * m;
  const T e = (i + 1 < L.nx) ? U(ku + 1, k + 1) : T(0);
  const T w = (i > 0) ? U(ku - 1, k - 1) : T(0);
  const T n = (j + 1 < L.ny) ? U(ku + ld, k + L.nx) : T(0);
  const T s = (j > 0) ? U(ku - ld, k - L.nx) : T(0);
  T lap;
  if constexpr (KIND == kDirichlet) {
    lap = (e + w - T(2) * c) * L.idx2 + (n + s - T(2) * c) * L.idy2;
  } else {
    const T cxe = L.cx[k], cxw = (i > 0) ? L.cx[k - 1] : T(0);
    const T cyn = L.cy[k], cys = (j > 0) ? L.cy[k - L.nx] : T(0);
    lap = (cxe * (e - c) - cxw * (c - w)) * L.idx2 + (cyn * (n - c) - cys * (c - s)) * L.idy2;
  }
  return (lap - L.lam * c) * m;
}

template <typename T, int KIND>
__device__ __forceinline__ T glob_inv(const LevelDev<T>& L, int64_t k) {
  if constexpr (KIND == kDirichlet) return L.inv_d * mval<T>(L.mask, k);
  return L.inv[k];
}

// ---------------------------------------------------------------------------
// Tiled smoother kernel.  MODE:
//   0 SMOOTH   : NS sweeps                                  -> u_out
//   1 DOWN     : NS sweeps, residual, restriction           -> u_out, C.b
//   2 UP       : u += P(C.u), NS sweeps                     -> u_out
//   3 UP_CG    : as UP with one more halo ring, then a = A z on the tile
//                interior -> u_out (z), az_out (a), and the nine PCG dots of
//                z, a, r (= b) and q_{k-1} (= w)
//   4 DOWN_STEP: (wavefront tiles only) the previous PCG step on the tile,
//                then DOWN from zero on the new residual -> upd.q_out,
//                upd.r_out, upd.p, upd.x, u_out, C.b
// ---------------------------------------------------------------------------
enum TileMode : int { kSmooth = 0, kDown = 1, kUp = 2, kUpCG = 3, kDownUpd = 4 };

// The PCG step that the level-0 down kernel of the next V-cycle applies
// first (kDownUpd; wavefront tiles only): with a = A z,
// q_out = a + beta q_in, r_out = r_in - alpha q_out, p = z + beta p,
// x += alpha p.  r_in and q_in are read on the tile's halo as well, so they
// cannot be the arrays written.
template <typename T>
struct PcgUpdArgs {
  const T* z;
  T* p;
  T* x;
  int64_t ld_x;
  const T* q_in;
  T* q_out;
  const T* r_in;  // passed to the kernel as its b
  T* r_out;
};

constexpr int kTx = 64, kTy = 16, kTileThreads = 256;

template <typename T>
struct TileArgs {
  LevelDev<T> L, C;
  const T* u_in;
  const T* b;
  T* u_out;
  const T* w;
  T* az_out;
  T omega;
  T omw[3];  // wave-tile kernels: the weight of sweep k (= omega unless the
             // plan smooths with Chebyshev weights)
  double* partials;
  PcgScalars* sc;
  int tiles_x, ntiles;
  PcgUpdArgs<T> upd;  // kDownUpd; b is r_in
};

template <typename T, int KIND>
__device__ __forceinline__ T prolong_at(const LevelDev<T>& C, int j, int i) {
  // _prolongate (multigrid.py:510-603) for fine cell (j, i), before x mask_f.
  const int J = j >> 1, I = i >> 1;
  const int sj = 2 * (j & 1) - 1, si = 2 * (i & 1) - 1;
  auto vm = [&](int jj, int ii, T& m) -> T {
    if (jj < 0 || jj >= C.ny || ii < 0 || ii >= C.nx) {
      m = T(0);
      return T(0);
    }
    const int64_t k = (int64_t)jj * C.nx + ii;
    m = mval<T>(C.mask, k);
    return C.u[k] * m;
  };
  T mc, mx, my, md;
  const T vc = vm(J, I, mc), vx = vm(J, I + si, mx), vy = vm(J + sj, I, my),
          vd = vm(J + sj, I + si, md);
  const T w = T(9) * vc + T(3) * vx + T(3) * vy + vd;
  // Dirichlet hierarchy: land and the space beyond the walls hold psi = 0, so
  // the weights are not renormalised (that would extrapolate to the wall).
  if constexpr (KIND == kDirichlet) return w * T(0.0625);
  const T d = T(9) * mc + T(3) * mx + T(3) * my + md;
  return w / (d > T(1) ? d : T(1));
}

// MK: the level has a land mask (staged in LDS); without one a cell's mask is
// just "inside the level", computed on the fly, which saves LDS (and raises
// occupancy) for all-wet levels.
template <typename T, int KIND, int NS, int MODE, bool ZERO, bool MK>
__device__ __forceinline__ void mg_tile_body(const TileArgs<T>& a, int bx) {
  // The convergence guard is checked after the halo loads are issued, so its
  // round trip overlaps theirs.
  const bool skip = a.sc && a.sc->done;
  constexpr int H = NS + ((MODE == kDown || MODE == kUpCG) ? 1 : 0);
  constexpr int W = kTx + 2 * H, Hh = kTy + 2 * H, N = W * Hh;
  __shared__ T sU[2][N];
  __shared__ T sB[N];
  __shared__ T sM[MK ? N : 1];
  __shared__ T sCX[KIND == kNoflux ? N : 1];
  __shared__ T sCY[KIND == kNoflux ? N : 1];
  __shared__ T sI[KIND == kNoflux ? N : 1];

  const LevelDev<T>& L = a.L;
  const int t = xcd_swizzle(bx, a.ntiles);
  const int i0 = (t % a.tiles_x) * kTx, j0 = (t / a.tiles_x) * kTy;
  const int tid = threadIdx.x;

  // Tiles whose halo lies inside the level need no range checks (all-wet
  // levels: their mask is 1 everywhere); uniform per block.
  const bool inner_tile = !MK && j0 - H >= 0 && j0 + kTy + H <= L.ny && i0 - H >= 0 &&
                          i0 + kTx + H <= L.nx;
  // ... and whose prolongation stencil stays inside the coarse level.
  const bool inner_coarse = !MK && j0 - H >= 1 && j0 + kTy + H <= L.ny - 1 && i0 - H >= 1 &&
                            i0 + kTx + H <= L.nx - 1;
  constexpr bool kUpMode = (MODE == kUp || MODE == kUpCG);
  constexpr int IT = (N + kTileThreads - 1) / kTileThreads;
  // Level l+1's iterate (times its mask) and mask around this tile for the
  // prolongation, staged in LDS: rows cj0 .. cj0 + CH - 1, columns
  // ci0 .. ci0 + CW - 1 of the coarse level (zero outside it).  With j0, i0
  // even the fine rows j0 - H .. j0 + kTy + H - 1 touch coarse rows
  // (j0 - H) / 2 - 1 .. (j0 + kTy + H - 1) / 2 + 1 (floor division).
  constexpr int CW = kTx / 2 + 2 * ((H + 1) / 2) + 2, CH = kTy / 2 + 2 * ((H + 1) / 2) + 2;
  // Staged in the spare sweep buffer sU[1]: the first sweep writes it only
  // after the barrier that ends the prolongation (keeps the level-0 tile's
  // LDS small enough for eight blocks per CU).
  static_assert(2 * CW * CH <= N, "coarse staging must fit in a sweep buffer");
  T* const sCu = sU[1];
  T* const sCm = sU[1] + CW * CH;
  const int cj0 = ((j0 - H) >> 1) - 1, ci0 = ((i0 - H) >> 1) - 1;
  T rU[kUpMode ? IT : 1], rM[kUpMode ? IT : 1];
  // q_{k-1} on the tile interior (kTx * kTy = 4 kTileThreads cells), loaded
  // up front so its round trip overlaps the sweeps.
  constexpr int kIn = kTx * kTy / kTileThreads;
  static_assert(kTx * kTy == kIn * kTileThreads, "tile interior must split evenly");
  T rQ[MODE == kUpCG ? kIn : 1];
  if constexpr (MODE == kUpCG) {
    const bool first = a.sc->first;
#pragma unroll
    for (int t = 0; t < kIn; ++t) {
      const int k = tid + t * kTileThreads;
      const int gj = j0 + k / kTx, gi = i0 + k % kTx;
      rQ[t] = (!first && gj < L.ny && gi < L.nx) ? a.w[gj * L.nx + gi] : T(0);
    }
  }

  // Fully unrolled (the trip count is a compile-time constant) so the
  // global loads of all iterations are in flight together; 32-bit offsets.
  // Tile coordinates advance incrementally (kTileThreads = dY W + dX).
  constexpr int dY = kTileThreads / W, dX = kTileThreads % W;
  int ty_ = tid / W, tx_ = tid % W;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int k = tid + it * kTileThreads;
    if (k >= N) break;
    const int y = ty_, x = tx_;
    tx_ += dX;
    ty_ += dY;
    if (tx_ >= W) {
      tx_ -= W;
      ++ty_;
    }
    const int gj = j0 - H + y, gi = i0 - H + x;
    const bool in = inner_tile || (gj >= 0 && gj < L.ny && gi >= 0 && gi < L.nx);
    const int g = in ? gj * L.nx + gi : 0;
    const T m = in ? mval<T>(L.mask, g) : T(0);
    if constexpr (MK) sM[k] = m;
    sB[k] = in ? a.b[g] : T(0);
    if constexpr (KIND == kNoflux) {
      sCX[k] = in ? L.cx[g] : T(0);
      sCY[k] = in ? L.cy[g] : T(0);
      sI[k] = in ? L.inv[g] : T(0);
    }
    T u = T(0);
    if constexpr (!ZERO) u = in ? a.u_in[g] * m : T(0);  // u_m = u * mask (multigrid.py:383)
    if constexpr (kUpMode) {
      rU[it] = u;
      rM[it] = m;
    } else {
      sU[0][k] = u;
    }
  }
  if constexpr (kUpMode) {
    const LevelDev<T>& C = a.C;
#pragma unroll
    for (int it = 0; it < (CW * CH + kTileThreads - 1) / kTileThreads; ++it) {
      const int k = tid + it * kTileThreads;
      if (k >= CW * CH) break;
      const int J = cj0 + k / CW, I = ci0 + k % CW;
      const bool in = (J >= 0 && J < C.ny && I >= 0 && I < C.nx);
      const int g = in ? J * C.nx + I : 0;
      const T m = in ? mval<T>(C.mask, g) : T(0);
      if constexpr (MK) sCm[k] = m;
      sCu[k] = in ? C.u[g] * m : T(0);
    }
    __syncthreads();
    // u = (u + P e_c) * mask  (multigrid.py:875-876), _prolongate
    // (multigrid.py:510-603) from the staged coarse tile.
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int k = tid + it * kTileThreads;
      if (k >= N) break;
      const int y = k / W, x = k % W;
      const int gj = j0 - H + y, gi = i0 - H + x;
      const int cy = (gj >> 1) - cj0, cx = (gi >> 1) - ci0;
      const int sj = (gj & 1) ? CW : -CW, si = (gi & 1) ? 1 : -1;
      const int q = cy * CW + cx;
      const T w = T(9) * sCu[q] + T(3) * sCu[q + si] + T(3) * sCu[q + sj] + sCu[q + sj + si];
      T d;
      if constexpr (MK) {
        d = T(9) * sCm[q] + T(3) * sCm[q + si] + T(3) * sCm[q + sj] + sCm[q + sj + si];
      } else if (inner_coarse) {  // all four coarse cells exist: weights sum to 16
        d = T(16);
      } else {  // coarse mask = inside the coarse level
        const int J = gj >> 1, I = gi >> 1, Jn = J + ((gj & 1) ? 1 : -1),
                  In = I + ((gi & 1) ? 1 : -1);
        const T mjc = (J >= 0 && J < a.C.ny) ? T(1) : T(0), mjn = (Jn >= 0 && Jn < a.C.ny) ? T(1) : T(0);
        const T mic = (I >= 0 && I < a.C.nx) ? T(1) : T(0), min_ = (In >= 0 && In < a.C.nx) ? T(1) : T(0);
        d = T(9) * mjc * mic + T(3) * mjc * min_ + T(3) * mjn * mic + mjn * min_;
      }
      const T dd = d > T(1) ? d : T(1);
      T e;
      if (KIND == kDirichlet || (!MK && inner_coarse)) e = w * T(0.0625);  // exact: w / 16
      else if constexpr (std::is_same<T, float>::value) e = w * __builtin_amdgcn_rcpf(dd);
      else e = w / dd;
      const T m = rM[it];
      sU[0][k] = (rU[it] + e * m) * m;
    }
  }
  if (skip) return;
  __syncthreads();

  // Mask of tile cell (y, x).
  auto Mq = [&](int y, int x) -> T {
    if constexpr (MK) {
      return sM[y * W + x];
    } else {
      const int gj = j0 - H + y, gi = i0 - H + x;
      return (gj >= 0 && gj < L.ny && gi >= 0 && gi < L.nx) ? T(1) : T(0);
    }
  };
  int cur = 0;
  auto sweep = [&](auto inner, int s) {
    const T om = a.omw[(s - 1) % 3];  // this sweep's weight
    // Valid region shrinks by one cell per sweep.
    const int lo = s, hiy = Hh - s, hix = W - s;
    const int ny_ = hiy - lo, nx_ = hix - lo;
    const int sdY = kTileThreads / nx_, sdX = kTileThreads % nx_;
    int sy = tid / nx_, sx = tid % nx_;
    for (int k = tid; k < ny_ * nx_; k += kTileThreads) {
      const int y = lo + sy, x = lo + sx;
      sx += sdX;
      sy += sdY;
      if (sx >= nx_) {
        sx -= nx_;
        ++sy;
      }
      const int q = y * W + x;
      T mq;
      if constexpr (decltype(inner)::value) mq = T(1);
      else mq = Mq(y, x);
      T inv;
      if constexpr (KIND == kDirichlet) inv = L.inv_d * mq;
      else inv = sI[q];
      T un;
      if (ZERO && s == 1 && !kUpMode) {
        un = (om * inv * sB[q]) * mq;
      } else {
        const T au = tile_Au<T, KIND, W>(sU[cur], sCX, sCY, mq, y, x, L.idx2, L.idy2, L.lam);
        un = (sU[cur][q] + om * inv * (sB[q] - au)) * mq;
      }
      sU[cur ^ 1][q] = un;
    }
  };
#pragma unroll
  for (int s = 1; s <= NS; ++s) {
    if (inner_tile) sweep(std::true_type{}, s);
    else sweep(std::false_type{}, s);
    __syncthreads();
    cur ^= 1;
  }

  if constexpr (MODE == kDown) {
    // r = (b - A u) * m on the tile interior, kept in the spare buffer.
    T* R = sU[cur ^ 1];
    for (int k = tid; k < kTx * kTy; k += kTileThreads) {
      const int y = H + k / kTx, x = H + k % kTx;
      const int q = y * W + x;
      const T mq = Mq(y, x);
      const T au = tile_Au<T, KIND, W>(sU[cur], sCX, sCY, mq, y, x, L.idx2, L.idy2, L.lam);
      R[q] = (sB[q] - au) * mq;
    }
    __syncthreads();
    // 2x2 mask-weighted restriction (multigrid.py:464-507).
    const LevelDev<T>& C = a.C;
    for (int k = tid; k < (kTx / 2) * (kTy / 2); k += kTileThreads) {
      const int cy_ = k / (kTx / 2), cx_ = k % (kTx / 2);
      const int J = j0 / 2 + cy_, I = i0 / 2 + cx_;
      if (J >= C.ny || I >= C.nx) continue;
      const int y = H + 2 * cy_, x = H + 2 * cx_;
      const int q00 = y * W + x, q10 = q00 + W, q01 = q00 + 1, q11 = q00 + W + 1;
      const T m00 = Mq(y, x), m10 = Mq(y + 1, x), m01 = Mq(y, x + 1), m11 = Mq(y + 1, x + 1);
      const T s4 = R[q00] * m00 + R[q10] * m10 + R[q01] * m01 + R[q11] * m11;
      const T cnt = m00 + m10 + m01 + m11;
      const int64_t g = (int64_t)J * C.nx + I;
      // Dirichlet: dry children count as zeros (psi on land), not renormalised
      const T rs = KIND == kDirichlet ? s4 * T(0.25) : s4 / (cnt > T(1) ? cnt : T(1));
      C.b[g] = rs * mval<T>(C.mask, g);
    }
  }

  if constexpr (MODE == kUpCG) {
    // z and a = A z on the tile interior (the sweeps left one valid ring
    // around it), and the PCG dots: summed in the level's precision within a
    // wave (FP32 on the FP32 hierarchy: its vectors carry FP32 rounding
    // already, and the scalars only steer the iteration -- the FP64
    // refinement certifies the solution), in FP64 across waves and blocks.
    T d[kNumCgDots];
#pragma unroll
    for (int k = 0; k < kNumCgDots; ++k) d[k] = T(0);
#pragma unroll
    for (int t = 0; t < kIn; ++t) {
      const int k = tid + t * kTileThreads;
      const int y = H + k / kTx, x = H + k % kTx;
      const int gj = j0 + k / kTx, gi = i0 + k % kTx;
      if (gj < L.ny && gi < L.nx) {
        const int q = y * W + x;
        const int g = gj * L.nx + gi;
        const T z = sU[cur][q];
        const T mq = inner_tile ? T(1) : Mq(y, x);
        const T az = tile_Au<T, KIND, W>(sU[cur], sCX, sCY, mq, y, x, L.idx2, L.idy2, L.lam);
        a.u_out[g] = z;
        a.az_out[g] = az;
        const T zd = z, rd = sB[q], ad = az, qd = rQ[t];
        d[kZR] += zd * rd;
        d[kZQ] += zd * qd;
        d[kZA] += zd * ad;
        d[kRR] += rd * rd;
        d[kRA] += rd * ad;
        d[kRQ] += rd * qd;
        d[kAA] += ad * ad;
        d[kAQ] += ad * qd;
        d[kQQ] += qd * qd;
      }
    }
    block_partials<kNumCgDots, kTileThreads>(d, a.partials, a.ntiles, bx);
  } else {
    // Write the tile interior.
#pragma unroll
    for (int k = tid; k < kTx * kTy; k += kTileThreads) {
      const int y = H + k / kTx, x = H + k % kTx;
      const int gj = j0 + k / kTx, gi = i0 + k % kTx;
      if (gj < L.ny && gi < L.nx) a.u_out[(int64_t)gj * L.nx + gi] = sU[cur][y * W + x];
    }
  }
}
template <typename T, int KIND, int NS, int MODE, bool ZERO, bool MK>
__global__ __launch_bounds__(kTileThreads) void mg_tile_kernel(TileArgs<T> a) {
  mg_tile_body<T, KIND, NS, MODE, ZERO, MK>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Wave-tile smoother kernel for FP32 levels with the Dirichlet operator
// (constant coefficients): the same modes and arithmetic as mg_tile_kernel,
// laid out for the wavefront.  One wavefront owns a tile of 64 columns x
// (ROWS + 2H) rows including its halo, each lane one column of u and b in
// registers.  East/west neighbours come from DPP wave shifts, north/south
// from the neighbouring registers, so the sweeps use no LDS, no barriers and
// no index arithmetic.  The tile interior is (64 - 2H) x ROWS; four
// independent tiles per 256-thread block.
//
// MK (land mask): each lane keeps one bit per row of its column (wet and
// inside the level).  Every iterate of a Dirichlet hierarchy is zero on land
// (updates and corrections are multiplied by the mask, coarse right-hand
// sides too), so u * mask is u: the stencils need no masked neighbours, and
// the mask enters as a select on each updated value, on A z and on the
// residual, and as the coarse mask on the restriction.  (A plain smoothing
// call with an arbitrary iterate keeps its land values in the reference; it
// stays on the LDS kernel.)
// ---------------------------------------------------------------------------

template <int NS, int MODE>
constexpr int wt_halo() {
  return NS + ((MODE == kDown || MODE == kDownUpd || MODE == kUpCG) ? 1 : 0);
}
// Lanes of halo on each side: one more for kDownUpd, whose a = A z is wrong
// on the wave's two edge lanes (their outer neighbours are in no lane).
template <int NS, int MODE>
constexpr int wt_lane_halo() {
  return wt_halo<NS, MODE>() + (MODE == kDownUpd ? 1 : 0);
}
// Interior rows per tile (ROWS; even, for the 2x2 restriction).  Short tiles
// keep the registers per lane low enough for 6-8 waves per SIMD, which hides
// the load latency better than the halo they save costs: measured at 2048^2
// with three fused sweeps, 8 rows (against 6, 10, 12, 16).  On small levels,
// whose kernels last about as long as one wave's chain of loads and
// arithmetic, tiles of 2 rows shorten that chain.
constexpr int kTileRows = 8, kShortTileRows = 2;
constexpr int64_t kShortTileMaxCells = 512 * 512;
// Rows of the level-0 PCG kernels (up leg with the dots, down leg with the
// PCG step).
template <int MODE>
constexpr int wt_pcg_rows() {
  return MODE == kUpCG ? 8 : 8;
}

template <int NS, int MODE, bool ZERO, int ROWS, bool MK, bool WALL = false>
__device__ __forceinline__ void mg_wave_tile_body(const TileArgs<float>& a, int bx, int nblk) {
  using T = float;
  constexpr int H = wt_halo<NS, MODE>();
  constexpr int HL = wt_lane_halo<NS, MODE>();
  constexpr int WTR = ROWS;
  constexpr int R = WTR + 2 * H;  // rows j0 - H .. j0 + WTR + H - 1
  constexpr bool kUpMode = (MODE == kUp || MODE == kUpCG);
  constexpr bool kDownMode = (MODE == kDown || MODE == kDownUpd);
  constexpr bool kUpd = MODE == kDownUpd;
  // Stop flag, checked once the loads are in flight (kDownUpd: 2 = apply the
  // PCG step, then stop).
  const int flag = a.sc ? a.sc->done : 0;
  const LevelDev<T>& L = a.L;
  const int lane = threadIdx.x & 63;
  // Tile of this wavefront, made scalar so that every per-tile quantity and
  // branch below is uniform (lane moves are only valid with all lanes on).
  const int tile = __builtin_amdgcn_readfirstlane(bx * 4 + (threadIdx.x >> 6));
  const bool active = tile < a.ntiles;
  const int tyi = active ? tile / a.tiles_x : 0;
  const int i0 = (active ? tile - tyi * a.tiles_x : 0) * (64 - 2 * HL);
  const int j0 = tyi * WTR;
  const int gi = i0 - HL + lane;
  const bool colin = active && gi >= 0 && gi < L.nx;
  // Interior columns of this lane, inside the level.
  const bool lin = colin && lane >= HL && lane < 64 - HL;
  const int I = gi >> 1, In = I + ((gi & 1) ? 1 : -1);
  // Coarse rows cr0 .. cr0 + NCR - 1 hold level l+1's iterate at this lane's
  // coarse column I = gi / 2 (cI) and at its neighbour In = I +- 1 on the
  // side of gi (cN), for every fine row of the tile.
  constexpr int RH = (-H) >> 1;
  constexpr int NCR = ((R - 1 - H) >> 1) - RH + 3;
  const int cr0 = (j0 >> 1) + RH - 1;
  // Cells outside the level are zero (the reference multiplies by the mask,
  // 0 outside and 1 inside an all-wet level).  Lanes outside are handled by
  // lane selects and zero update coefficients; rows outside exist only in the
  // tiles at the top and bottom of the level (wave-uniform), which zero them
  // in a branch of their own, so that no other tile pays per-row conditions.
  constexpr int HZ = kUpd ? 1 : 0;  // rows of z beyond the tile (kDownUpd)
  const bool edge = j0 - H - HZ < 0 || j0 - H + R + HZ > L.ny ||
                    (kUpMode && (cr0 < 0 || cr0 + NCR > a.C.ny));
  auto row_in = [&](int r) -> bool {
    const int gj = j0 - H + r;
    return gj >= 0 && gj < L.ny;
  };

  // Loads from clamped (always valid) cells, 32-bit offsets.
  const unsigned gic = (unsigned)min(max(gi, 0), L.nx - 1);
  const int nL = L.ny * L.nx;
  T u[R], b[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int gjc = min(max(j0 - H + r, 0), L.ny - 1);
    const T bv = buf_load(buf_rsrc(a.b, nL), gic, gjc * L.nx);
    b[r] = colin ? bv : T(0);
    if constexpr (ZERO) {
      u[r] = T(0);
    } else {
      const T uv = buf_load(buf_rsrc(a.u_in, nL), gic, gjc * L.nx);
      u[r] = colin ? uv : T(0);  // u * mask (multigrid.py:383)
    }
  }
  // MK: bit r = row r of this lane's column is wet and inside the level
  // (rows outside the level are cleared with the edge rows below).
  uint32_t mb = 0;
  if constexpr (MK) {
    static_assert(R <= 32, "one 32-bit mask word per lane");
    if (L.mcol) {
      // rows g .. g + 31 from two words of the column-packed mask (rows above
      // the level, g0 < 0, take row 0's bits shifted in: they are cleared
      // with the edge rows below, as are rows below the level)
      const int g0 = j0 - H, g = max(g0, 0);
      const int blk = g >> 5;
      const uint64_t w0 = L.mcol[blk * L.nx + (int)gic];
      const uint64_t w1 = L.mcol[(blk + 1) * L.nx + (int)gic];
      const uint64_t win = ((w1 << 32) | w0) >> (g & 31);
      mb = (uint32_t)(win << (g - g0));
      if (R < 32) mb &= (1u << R) - 1u;
    } else {
      mb = (R >= 32) ? ~0u : ((1u << R) - 1u);
    }
    if (!colin) mb = 0;
  }
  auto wet_r = [&](int r) -> bool { return MK ? ((mb >> r) & 1u) != 0 : colin; };
  const bool first = (MODE == kUpCG) ? a.sc->first : true;
  // (Each lane loads both coarse columns: taking the neighbour column from
  // the lane beside it, possible for odd H, held more registers than the
  // loads it saved.)
  T cI[kUpMode ? NCR : 1], cN[kUpMode ? NCR : 1];
  if constexpr (kUpMode) {
    const LevelDev<T>& C = a.C;
    const bool ci = active && I >= 0 && I < C.nx, cn = active && In >= 0 && In < C.nx;
    const unsigned Ic = (unsigned)min(max(I, 0), C.nx - 1);
    const unsigned Inc = (unsigned)min(max(In, 0), C.nx - 1);
#pragma unroll
    for (int p = 0; p < NCR; ++p) {
      const int crow = min(max(cr0 + p, 0), C.ny - 1) * C.nx;
      const T vi = buf_load(buf_rsrc(C.u, C.ny * C.nx), Ic, crow);
      const T vn = buf_load(buf_rsrc(C.u, C.ny * C.nx), Inc, crow);
      cI[p] = ci ? vi : T(0);
      cN[p] = cn ? vn : T(0);
    }
  }
  // kDownUpd: z on rows -1 .. R of the tile, q_in on its rows.
  T zr[kUpd ? R + 2 : 1], qv[kUpd ? R : 1];
  if constexpr (kUpd) {
#pragma unroll
    for (int t = 0; t < R + 2; ++t) {
      const int gjc = min(max(j0 - H - 1 + t, 0), L.ny - 1);
      const T v = buf_load(buf_rsrc(a.upd.z, nL), gic, gjc * L.nx);
      zr[t] = colin ? v : T(0);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int gjc = min(max(j0 - H + r, 0), L.ny - 1);
      const T v = buf_load(buf_rsrc(a.upd.q_in, nL), gic, gjc * L.nx);
      qv[r] = colin ? v : T(0);
    }
  }
  if (kUpd ? flag == 1 : flag != 0) return;
  // Row factor of the prolongation weights (below): 1/4 where both coarse
  // rows exist, 0 on rows outside the level.
  T rowf[kUpMode ? R : 1];
#pragma unroll
  for (int r = 0; r < (kUpMode ? R : 1); ++r) rowf[r] = T(0.25);
  if (edge) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!row_in(r)) {
        u[r] = T(0);
        b[r] = T(0);
        if constexpr (kUpd) qv[r] = T(0);
        if constexpr (MK) mb &= ~(1u << r);
      }
    }
    if constexpr (kUpd) {
#pragma unroll
      for (int t = 0; t < R + 2; ++t) {
        const int gj = j0 - H - 1 + t;
        if (gj < 0 || gj >= L.ny) zr[t] = T(0);
      }
    }
    if constexpr (kUpMode) {
      const LevelDev<T>& C = a.C;
#pragma unroll
      for (int p = 0; p < NCR; ++p) {
        const int J = cr0 + p;
        if (J < 0 || J >= C.ny) {
          cI[p] = T(0);
          cN[p] = T(0);
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int gj = j0 - H + r;
        const int J = gj >> 1, Jn = J + ((gj & 1) ? 1 : -1);
        // Dirichlet weights are not renormalised at the walls (the coarse
        // values beyond them are zero): 1/4 on every row inside the level.
        (void)J;
        (void)Jn;
        rowf[r] = row_in(r) ? T(0.25) : T(0);
      }
    }
  }
  if constexpr (MK && !ZERO) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (!wet_r(r)) u[r] = T(0);  // u * mask
  }
  // Rows outside the level back to zero after a sweep (edge tiles only).
  auto zero_outer_rows = [&]() {
    if (edge) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (!row_in(r)) u[r] = T(0);
    }
  };

  const T idx2 = L.idx2, idy2 = L.idy2, lam = L.lam, invd = L.inv_d;
  // WALL: per row of this lane's column, the diagonal correction
  // wall_g (dry x neighbours / dx^2 + dry y neighbours / dy^2) and the
  // cell's inverse diagonal (0 on dry or outside cells).  Neighbours outside
  // the tile's rows or lanes only occur in its halo, whose values the sweeps
  // discard.
  T dgr[WALL ? R : 1], ivr[WALL ? R : 1];
  if constexpr (WALL) {
    uint32_t wm = mb;
    if constexpr (!MK) {
      wm = colin ? ((R >= 32) ? ~0u : ((1u << R) - 1u)) : 0u;
      if (edge) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (!row_in(r)) wm &= ~(1u << r);
      }
    }
    const uint32_t wb = (uint32_t)__builtin_amdgcn_mov_dpp((int)wm, 0x138, 0xf, 0xf, true);
    const uint32_t eb = (uint32_t)__builtin_amdgcn_mov_dpp((int)wm, 0x130, 0xf, 0xf, true);
    const T g = L.wall_g, d0 = -(T(2) * idx2 + T(2) * idy2 + lam);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const T nlx = T(2 - (int)((wb >> r) & 1u) - (int)((eb >> r) & 1u));
      const T nly = T(2 - (r + 1 < R ? (int)((wm >> (r + 1)) & 1u) : 1) -
                      (r > 0 ? (int)((wm >> (r - 1)) & 1u) : 1));
      const T dg = g * (nlx * idx2 + nly * idy2);
      dgr[r] = dg;
      ivr[r] = ((wm >> r) & 1u) ? T(1) / (d0 + dg) : T(0);
    }
  }
  if constexpr (kUpd) {
    // The previous PCG step (pcg_cg_update_kernel's arithmetic) on every
    // cell of the tile, so that the sweeps below start from r_{k+1}:
    // a = A z, q_k = a + beta q_{k-1}, r_{k+1} = r_k - alpha q_k; beta == 0
    // (the first step or a restart) does not read p_{k-1}, q_{k-1}.
    const double beta64 = a.sc->beta;
    const T alpha = (T)a.sc->alpha, beta = (T)beta64;
    const bool restart = beta64 == 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const T c = zr[r + 1];
      const T e = lane_east(c), w = lane_west(c);
      T av = (e + w - T(2) * c) * idx2 + (zr[r + 2] + zr[r] - T(2) * c) * idy2 - lam * c;
      if constexpr (MK) av = wet_r(r) ? av : T(0);  // A z = (lap z - lam z) m
      const T qn = restart ? av : av + beta * qv[r];
      qv[r] = qn;
      b[r] = colin ? b[r] - alpha * qn : T(0);  // zero outside the level
    }
    if (edge) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (!row_in(r)) b[r] = T(0);
    }
    // Store q_k and r_{k+1}; p_k = z + beta p_{k-1}, x += alpha p_k, on the
    // tile interior (in place: no other tile reads p or x).
    if (lin) {
#pragma unroll
      for (int t = 0; t < WTR; ++t) {
        const int gj = j0 + t;
        if (gj < L.ny) {
          const int r = H + t;
          buf_store(buf_rsrc(a.upd.q_out, nL), (unsigned)gi, gj * L.nx, qv[r]);
          buf_store(buf_rsrc(a.upd.r_out, nL), (unsigned)gi, gj * L.nx, b[r]);
          const int64_t k = (int64_t)gj * L.nx + gi;
          const T zc = zr[r + 1];
          const T pn = restart ? zc : zc + beta * a.upd.p[k];
          a.upd.p[k] = pn;
          T* xp = a.upd.x + (int64_t)gj * a.upd.ld_x + gi;
          *xp = *xp + alpha * pn;
        }
      }
    }
    if (flag != 0) return;  // flag 2: that was the last step
  }
  // (A u)(r) for the tile's current u, inside the level.
  auto Au = [&](int r) -> T {
    const T c = u[r];
    const T e = lane_east(c), w = lane_west(c);
    const T lap = (e + w - T(2) * c) * idx2 + (u[r + 1] + u[r - 1] - T(2) * c) * idy2;
    if constexpr (WALL) return lap - lam * c + dgr[r] * c;
    return lap - lam * c;
  };
  // u = (u + P e_c) * mask  (multigrid.py:875-876), _prolongate
  // (multigrid.py:510-603).  The weight normaliser is separable: with the
  // coarse row J = gj / 2, its neighbour Jn on the side of gj, and the same
  // for columns, d = 9 mJ mI + 3 mJ mIn + 3 mJn mI + mJn mIn
  //                = (3 mJ + mJn) (3 mI + mIn),
  // 16 wherever all four coarse cells exist; where the row or column factor
  // is zero, so is every weighted coarse value.
  if constexpr (kUpMode) {
    const T lanef = colin ? T(0.25) : T(0);  // as rowf: no renormalisation at the walls
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int p = ((r - H) >> 1) - RH + 1;
      const int pn = p + (((r + H) & 1) ? 1 : -1);
      const T w = T(9) * cI[p] + T(3) * cN[p] + T(3) * cI[pn] + cN[pn];
      const T un = u[r] + w * (rowf[r] * lanef);
      if constexpr (MK) u[r] = wet_r(r) ? un : T(0);  // (u + P e) * mask
      else u[r] = un;
    }
  }
  // Damped-Jacobi sweeps (multigrid.py:408-461); the valid region shrinks
  // by one row and one lane per sweep.  prev keeps row r - 1's old value.
  // u + om (b - A u) / diag with A u = idx2 (e + w) + idy2 (n + s) - (2 idx2
  // + 2 idy2 + lam) u and 1 / diag = invd, one coefficient per term; the
  // coefficients are zero on lanes outside the level.
  const T lk = colin ? T(1) : T(0);
#pragma unroll
  for (int sw = 1; sw <= NS; ++sw) {
    // this sweep's weight (uniform: omega, or a Chebyshev weight)
    const T oms = a.omw[(sw - 1) % 3];
    const T kb = oms * invd * lk, kx = -oms * invd * idx2 * lk, ky = -oms * invd * idy2 * lk;
    const T k0 = (T(1) + oms * invd * (T(2) * idx2 + T(2) * idy2 + lam)) * lk;
    T prev = u[sw - 1];
#pragma unroll
    for (int r = sw; r < R - sw; ++r) {
      const T c = u[r];
      T un;
      if constexpr (WALL) {
        // u + om (b - A u) / d = (1 - om) u + om (b - offdiag(u)) / d with
        // the cell's own d
        if (ZERO && sw == 1 && !kUpMode) {
          un = (oms * ivr[r]) * b[r];
        } else {
          const T e = lane_east(c), w = lane_west(c);
          un = (T(1) - oms) * c +
               (oms * ivr[r]) * (b[r] - idx2 * (e + w) - idy2 * (u[r + 1] + prev));
        }
      } else if (ZERO && sw == 1 && !kUpMode) {
        un = kb * b[r];
      } else {
        const T e = lane_east(c), w = lane_west(c);
        un = k0 * c + kb * b[r] + kx * (e + w) + ky * (u[r + 1] + prev);
      }
      prev = c;
      if constexpr (MK) u[r] = wet_r(r) ? un : T(0);
      else u[r] = un;
    }
    zero_outer_rows();
  }

  // PCG dots (UP_CG), reduced at the end.
  T d[kNumCgDots];
#pragma unroll
  for (int k = 0; k < kNumCgDots; ++k) d[k] = T(0);
  if constexpr (kDownMode) {
    // r = (b - A u) * m on the interior rows, 2x2 restriction (sum / 4):
    // rows summed in registers, the column pair
    // (gi even, gi + 1) through one lane shift.
    const LevelDev<T>& C = a.C;
    const bool even = lin && ((gi & 1) == 0);
    // MK: bit k2 = the coarse mask of coarse row j0 / 2 + k2 in column I,
    // from two words of the coarse level's column-packed mask
    uint32_t cmb = ~0u;
    if constexpr (MK) {
      static_assert(WTR / 2 <= 32, "one 32-bit coarse mask word per lane");
      if (C.mcol) {
        const int cg = j0 >> 1, blk = cg >> 5;
        const int Icl = min(max(I, 0), C.nx - 1);
        const uint64_t w0 = C.mcol[blk * C.nx + Icl];
        const uint64_t w1 = C.mcol[(blk + 1) * C.nx + Icl];
        cmb = (uint32_t)(((w1 << 32) | w0) >> (cg & 31));
      }
    }
#pragma unroll
    for (int k2 = 0; k2 < WTR / 2; ++k2) {
      const int r0 = H + 2 * k2;
      const bool w0 = MK ? wet_r(r0) : (colin && row_in(r0));
      const bool w1 = MK ? wet_r(r0 + 1) : (colin && row_in(r0 + 1));
      // (A u with every lane active: lane moves read the neighbouring lanes)
      const T a0 = Au(r0), a1 = Au(r0 + 1);
      const T res0 = w0 ? b[r0] - a0 : T(0), res1 = w1 ? b[r0 + 1] - a1 : T(0);
      const T s2 = res0 + res1;
      const T s4 = s2 + lane_east(s2);
      const int J = (j0 >> 1) + k2;
      if (even && J < C.ny && I < C.nx) {
        // children beyond the level or on land count as zeros (the Dirichlet
        // restriction); the coarse level of an all-wet level is all wet
        T v = s4 * T(0.25);
        if constexpr (MK)
          if (!((cmb >> k2) & 1u)) v = T(0);
        buf_store(buf_rsrc(C.b, C.ny * C.nx), (unsigned)I, J * C.nx, v);
      }
    }
  }
  if constexpr (MODE == kUpCG) {
    // z, a = A z on the interior and the PCG dots (see mg_tile_kernel).
    // q_{k-1} is read only here (a compiler barrier keeps its loads from
    // being hoisted to the kernel start, where u, b and the coarse rows
    // already fill the registers).
    asm volatile("" ::: "memory");
    T rq[WTR];
#pragma unroll
    for (int t = 0; t < WTR; ++t) {
      const T v = buf_load(buf_rsrc(a.w, nL), gic, min(j0 + t, L.ny - 1) * L.nx);
      rq[t] = (!first && lin) ? v : T(0);
    }
    if (edge) {
#pragma unroll
      for (int t = 0; t < WTR; ++t)
        if (!row_in(H + t)) rq[t] = T(0);
    }
    // Every product below has a factor that is zero outside the tile
    // interior (rd, qd, or az): z itself is only stored.  (A z is formed
    // with every lane active: lane moves read the neighbouring lanes.)
    T azs[WTR];
#pragma unroll
    for (int t = 0; t < WTR; ++t) {
      const int r = H + t;
      const T azv = Au(r);
      const T az = (lin && (MK ? wet_r(r) : (!edge || row_in(r)))) ? azv : T(0);
      azs[t] = az;
      const T z = u[r], rd = lin ? b[r] : T(0), qd = rq[t];
      d[kZR] += z * rd;
      d[kZQ] += z * qd;
      d[kZA] += z * az;
      d[kRR] += rd * rd;
      d[kRA] += rd * az;
      d[kRQ] += rd * qd;
      d[kAA] += az * az;
      d[kAQ] += az * qd;
      d[kQQ] += qd * qd;
    }
    // z; a = A z too when az_out is given (batched solves), else the PCG
    // update kernel forms A z again from z.
    if (lin) {
#pragma unroll
      for (int t = 0; t < WTR; ++t) {
        const int gj = j0 + t;
        if (gj < L.ny) buf_store(buf_rsrc(a.u_out, nL), (unsigned)gi, gj * L.nx, u[H + t]);
      }
      if (a.az_out) {
#pragma unroll
        for (int t = 0; t < WTR; ++t) {
          const int gj = j0 + t;
          if (gj < L.ny) buf_store(buf_rsrc(a.az_out, nL), (unsigned)gi, gj * L.nx, azs[t]);
        }
      }
    }
  } else {
    if (lin) {
#pragma unroll
      for (int t = 0; t < WTR; ++t) {
        const int gj = j0 + t;
        if (gj < L.ny) buf_store(buf_rsrc(a.u_out, nL), (unsigned)gi, gj * L.nx, u[H + t]);
      }
    }
  }
  if constexpr (MODE == kUpCG) block_partials<kNumCgDots, 256>(d, a.partials, nblk, bx);
}
template <int NS, int MODE, bool ZERO, int ROWS, bool MK = false, bool WALL = false>
__global__ __launch_bounds__(256) void mg_wave_tile_kernel(TileArgs<float> a) {
  mg_wave_tile_body<NS, MODE, ZERO, ROWS, MK, WALL>(a, xcd_swizzle(blockIdx.x, gridDim.x),
                                                    gridDim.x);
}

// ---------------------------------------------------------------------------
// Single-workgroup coarse V-cycle (levels lc .. nL-1) on global arrays that
// stay resident in L2.  Starts from a zero guess at level lc (preconditioner
// use) and writes the result to levels[lc].u.
// ---------------------------------------------------------------------------
constexpr int kCoarseThreads = 1024;
// Per-sweep weights the coarse kernel carries for each leg (sweeps past the
// last take the last weight).
constexpr int kMaxW = 8;

template <typename T>
struct CoarseArgs {
  LevelDev<T> lv[16];
  int lc, nl, n_pre, n_post, n_coarse;
  int nonzero_init;  // level lc starts from lv[lc].u instead of zero
  T omega;
  T w[3][kMaxW];  // sweep weights: pre-smoothing, post-smoothing, bottom
  const int* done;
};

template <typename T, int KIND>
__device__ void cw_sweeps(const LevelDev<T>& L, int n, bool zero, const T* w) {
  const int N = L.ny * L.nx;
  T* src = L.u;
  T* dst = L.t;
  for (int s = 0; s < n; ++s) {
    const T om = w[s < kMaxW ? s : kMaxW - 1];
    for (int k = threadIdx.x; k < N; k += kCoarseThreads) {
      const int j = k / L.nx, i = k % L.nx;
      const T m = mval<T>(L.mask, k);
      const T inv = glob_inv<T, KIND>(L, k);
      T un;
      if (zero && s == 0) un = (om * inv * L.b[k]) * m;
      else un = (src[k] + om * inv * (L.b[k] - glob_Au<T, KIND>(L, src, j, i))) * m;
      dst[k] = un;
    }
    __syncthreads();
    T* tmp = src;
    src = dst;
    dst = tmp;
  }
  if (src != L.u) {  // odd number of sweeps: copy back
    for (int k = threadIdx.x; k < N; k += kCoarseThreads) L.u[k] = src[k];
    __syncthreads();
  }
}

// Level l of the coarse kernel with its iterate, scratch and right-hand side
// in LDS (u, t, b at cell offset off), coefficients and masks in global memory.
template <typename T, bool LDS>
__device__ __forceinline__ LevelDev<T> lds_level(const CoarseArgs<T>& a, int l, T* sm, int off) {
  LevelDev<T> L = a.lv[l];
  if constexpr (!LDS) return L;
  const int n = L.ny * L.nx;
  L.u = sm + off;
  L.t = sm + off + n;
  L.b = sm + off + 2 * n;
  return L;
}

// LDS = false keeps every level in global memory (bottom levels too large
// for LDS: grids that stop halving early).
template <typename T, int KIND, bool LDS>
__global__ __launch_bounds__(kCoarseThreads) void mg_coarse_kernel(CoarseArgs<T> a) {
  const bool skip = a.done && *a.done;
  if (!LDS && skip) return;
  extern __shared__ __align__(16) unsigned char coarse_smem[];
  T* sm = reinterpret_cast<T*>(coarse_smem);
  int off[16];
  {
    int o = 0;
    for (int l = a.lc; l < a.nl; ++l) {
      off[l] = o;
      o += 3 * a.lv[l].ny * a.lv[l].nx;
    }
  }
  // Level lc's right-hand side (and guess) from global memory.
  if constexpr (LDS) {
    const LevelDev<T>& G = a.lv[a.lc];
    const LevelDev<T> L = lds_level<T, LDS>(a, a.lc, sm, off[a.lc]);
    for (int k = threadIdx.x; k < G.ny * G.nx; k += kCoarseThreads) {
      L.b[k] = G.b[k];
      if (a.nonzero_init) L.u[k] = G.u[k];
    }
    if (skip) return;
    __syncthreads();
  }
  // Down leg.
  for (int l = a.lc; l < a.nl - 1; ++l) {
    const LevelDev<T> L = lds_level<T, LDS>(a, l, sm, off[l]);
    const LevelDev<T> C = lds_level<T, LDS>(a, l + 1, sm, off[l + 1]);
    const bool zero = !(a.nonzero_init && l == a.lc);
    if (a.n_pre > 0) {
      cw_sweeps<T, KIND>(L, a.n_pre, zero, a.w[0]);
    } else if (zero) {
      for (int k = threadIdx.x; k < L.ny * L.nx; k += kCoarseThreads) L.u[k] = T(0);
      __syncthreads();
    }
    for (int k = threadIdx.x; k < C.ny * C.nx; k += kCoarseThreads) {
      const int J = k / C.nx, I = k % C.nx;
      T s4 = T(0), cnt = T(0);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int dj = (q == 1 || q == 3), di = (q >= 2);
        const int j = 2 * J + dj, i = 2 * I + di;
        const int64_t f = (int64_t)j * L.nx + i;
        const T m = mval<T>(L.mask, f);
        const T r = (L.b[f] - glob_Au<T, KIND>(L, L.u, j, i)) * m;
        s4 += r * m;
        cnt += m;
      }
      const T rs = KIND == kDirichlet ? s4 * T(0.25) : s4 / (cnt > T(1) ? cnt : T(1));
      C.b[k] = rs * mval<T>(C.mask, k);
    }
    __syncthreads();
  }
  // Bottom solve: n_coarse sweeps from zero (from the guess when the whole
  // hierarchy is one level and a guess was given).
  {
    const LevelDev<T> B = lds_level<T, LDS>(a, a.nl - 1, sm, off[a.nl - 1]);
    const bool zero = !(a.nonzero_init && a.lc == a.nl - 1);
    if (a.n_coarse > 0) {
      cw_sweeps<T, KIND>(B, a.n_coarse, zero, a.w[2]);
    } else if (zero) {
      for (int k = threadIdx.x; k < B.ny * B.nx; k += kCoarseThreads) B.u[k] = T(0);
      __syncthreads();
    }
  }
  // Up leg.
  for (int l = a.nl - 2; l >= a.lc; --l) {
    const LevelDev<T> L = lds_level<T, LDS>(a, l, sm, off[l]);
    const LevelDev<T> C = lds_level<T, LDS>(a, l + 1, sm, off[l + 1]);
    for (int k = threadIdx.x; k < L.ny * L.nx; k += kCoarseThreads) {
      const int j = k / L.nx, i = k % L.nx;
      const T m = mval<T>(L.mask, k);
      L.u[k] = (L.u[k] + prolong_at<T, KIND>(C, j, i) * m) * m;
    }
    __syncthreads();
    cw_sweeps<T, KIND>(L, a.n_post, false, a.w[1]);
  }
  // Level lc's corrected iterate back to global memory.
  if constexpr (LDS) {
    const LevelDev<T>& G = a.lv[a.lc];
    const LevelDev<T> L = lds_level<T, LDS>(a, a.lc, sm, off[a.lc]);
    for (int k = threadIdx.x; k < G.ny * G.nx; k += kCoarseThreads) G.u[k] = L.u[k];
  }
}

// LDS-resident coarse V-cycle.  Levels lc .. nl-1 keep their iterate u,
// scratch t and right-hand side b in dynamic LDS, addressed by cell offsets
// (ds_read/ds_write); the 1024 threads tile each level as 32 x 32 blocks of
// cells, so no integer division is needed.  Every stored iterate is already
// multiplied by its mask, which is the reference's u * mask before the
// stencil (multigrid.py:383).
template <typename T, int KIND>
struct CoarseLds {
  T* sm;
  __device__ __forceinline__ T m_at(const LevelDev<T>& L, int k) const {
    return mval<T>(L.mask, k);
  }
  __device__ __forceinline__ T inv_at(const LevelDev<T>& L, int k, T m) const {
    if constexpr (KIND == kDirichlet) return L.inv_d * m;
    return L.inv[k];
  }
  // (A u)(j, i) for u at LDS offset ou (zero outside the level).
  __device__ __forceinline__ T Au(const LevelDev<T>& L, int ou, int j, int i, int k, T m) const {
    const T c = sm[ou + k];
    const T e = (i + 1 < L.nx) ? sm[ou + k + 1] : T(0);
    const T w = (i > 0) ? sm[ou + k - 1] : T(0);
    const T n = (j + 1 < L.ny) ? sm[ou + k + L.nx] : T(0);
    const T s = (j > 0) ? sm[ou + k - L.nx] : T(0);
    T lap;
    if constexpr (KIND == kDirichlet) {
      lap = (e + w - T(2) * c) * L.idx2 + (n + s - T(2) * c) * L.idy2;
    } else {
      const T cxe = L.cx[k], cxw = (i > 0) ? L.cx[k - 1] : T(0);
      const T cyn = L.cy[k], cys = (j > 0) ? L.cy[k - L.nx] : T(0);
      lap = (cxe * (e - c) - cxw * (c - w)) * L.idx2 + (cyn * (n - c) - cys * (c - s)) * L.idy2;
    }
    return (lap - L.lam * c) * m;
  }
  // n damped-Jacobi sweeps of level L (u at ou, scratch at ot, b at ob); the
  // result ends in ou.
  __device__ void sweeps(const LevelDev<T>& L, int ou, int ot, int ob, int n, bool zero,
                         const T* w, int tx, int ty) const {
    int src = ou, dst = ot;
    for (int sw = 0; sw < n; ++sw) {
      const T om = w[sw < kMaxW ? sw : kMaxW - 1];
      for (int j = ty; j < L.ny; j += 32)
        for (int i = tx; i < L.nx; i += 32) {
          const int k = j * L.nx + i;
          const T m = m_at(L, k);
          const T inv = inv_at(L, k, m);
          T un;
          if (zero && sw == 0) un = (om * inv * sm[ob + k]) * m;
          else un = (sm[src + k] + om * inv * (sm[ob + k] - Au(L, src, j, i, k, m))) * m;
          sm[dst + k] = un;
        }
      __syncthreads();
      const int tmp = src;
      src = dst;
      dst = tmp;
    }
    if (src != ou) {  // odd number of sweeps: copy back
      for (int j = ty; j < L.ny; j += 32)
        for (int i = tx; i < L.nx; i += 32) sm[ou + j * L.nx + i] = sm[src + j * L.nx + i];
      __syncthreads();
    }
  }
};

template <typename T, int KIND>
__device__ __forceinline__ void mg_coarse_lds_body(const CoarseArgs<T>& a,
                                                   unsigned char* coarse_smem) {
  const bool skip = a.done && *a.done;
  CoarseLds<T, KIND> S{reinterpret_cast<T*>(coarse_smem)};
  T* sm = S.sm;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  // LDS offsets of each level's u, t, b (a table in LDS: indexed by the
  // runtime level, a private array would live in scratch memory).
  __shared__ int ou[16], ot[16], ob[16];
  if (threadIdx.x == 0) {
    int o = 0;
    for (int l = a.lc; l < a.nl; ++l) {
      const int n = a.lv[l].ny * a.lv[l].nx;
      ou[l] = o;
      ot[l] = o + n;
      ob[l] = o + 2 * n;
      o += 3 * n;
    }
  }
  __syncthreads();
  // Level lc's right-hand side (and masked guess) from global memory.
  {
    const LevelDev<T>& G = a.lv[a.lc];
    for (int j = ty; j < G.ny; j += 32)
      for (int i = tx; i < G.nx; i += 32) {
        const int k = j * G.nx + i;
        sm[ob[a.lc] + k] = G.b[k];
        if (a.nonzero_init) sm[ou[a.lc] + k] = G.u[k] * S.m_at(G, k);
      }
    if (skip) return;
    __syncthreads();
  }
  // Down leg: pre-smoothing, residual, mask-weighted 2x2 restriction.
  for (int l = a.lc; l < a.nl - 1; ++l) {
    const LevelDev<T>& L = a.lv[l];
    const LevelDev<T>& C = a.lv[l + 1];
    const bool zero = !(a.nonzero_init && l == a.lc);
    if (a.n_pre > 0) {
      S.sweeps(L, ou[l], ot[l], ob[l], a.n_pre, zero, a.w[0], tx, ty);
    } else if (zero) {
      for (int j = ty; j < L.ny; j += 32)
        for (int i = tx; i < L.nx; i += 32) sm[ou[l] + j * L.nx + i] = T(0);
      __syncthreads();
    }
    // residual r = (b - A u) m into the scratch array
    for (int j = ty; j < L.ny; j += 32)
      for (int i = tx; i < L.nx; i += 32) {
        const int k = j * L.nx + i;
        const T m = S.m_at(L, k);
        sm[ot[l] + k] = (sm[ob[l] + k] - S.Au(L, ou[l], j, i, k, m)) * m;
      }
    __syncthreads();
    for (int J = ty; J < C.ny; J += 32)
      for (int I = tx; I < C.nx; I += 32) {
        const int f = (2 * J) * L.nx + 2 * I;
        const T m00 = S.m_at(L, f), m01 = S.m_at(L, f + 1), m10 = S.m_at(L, f + L.nx),
                m11 = S.m_at(L, f + L.nx + 1);
        const T s4 = sm[ot[l] + f] * m00 + sm[ot[l] + f + 1] * m01 + sm[ot[l] + f + L.nx] * m10 +
                     sm[ot[l] + f + L.nx + 1] * m11;
        const T cnt = m00 + m01 + m10 + m11;
        const int kc = J * C.nx + I;
        const T rs = KIND == kDirichlet ? s4 * T(0.25) : s4 / (cnt > T(1) ? cnt : T(1));
        sm[ob[l + 1] + kc] = rs * S.m_at(C, kc);
      }
    __syncthreads();
  }
  // Bottom solve.
  {
    const int lb = a.nl - 1;
    const LevelDev<T>& B = a.lv[lb];
    const bool zero = !(a.nonzero_init && a.lc == lb);
    if (a.n_coarse > 0) {
      S.sweeps(B, ou[lb], ot[lb], ob[lb], a.n_coarse, zero, a.w[2], tx, ty);
    } else if (zero) {
      for (int j = ty; j < B.ny; j += 32)
        for (int i = tx; i < B.nx; i += 32) sm[ou[lb] + j * B.nx + i] = T(0);
      __syncthreads();
    }
  }
  // Up leg: u += P(u_c) (9/3/3/1 weights, mask renormalised), post-smoothing.
  for (int l = a.nl - 2; l >= a.lc; --l) {
    const LevelDev<T>& L = a.lv[l];
    const LevelDev<T>& C = a.lv[l + 1];
    for (int j = ty; j < L.ny; j += 32)
      for (int i = tx; i < L.nx; i += 32) {
        const int k = j * L.nx + i;
        const T m = S.m_at(L, k);
        const int J = j >> 1, I = i >> 1;
        const int sj = (j & 1) ? 1 : -1, si = (i & 1) ? 1 : -1;
        auto vm = [&](int jj, int ii, T& mm) -> T {
          if (jj < 0 || jj >= C.ny || ii < 0 || ii >= C.nx) {
            mm = T(0);
            return T(0);
          }
          const int kc = jj * C.nx + ii;
          mm = S.m_at(C, kc);
          return sm[ou[l + 1] + kc];  // stored iterates are already masked
        };
        T mc, mx, my, md;
        const T vc = vm(J, I, mc), vx = vm(J, I + si, mx), vy = vm(J + sj, I, my),
                vd = vm(J + sj, I + si, md);
        const T w = T(9) * vc + T(3) * vx + T(3) * vy + vd;
        const T d = T(9) * mc + T(3) * mx + T(3) * my + md;
        const T e = KIND == kDirichlet ? w * T(0.0625) : w / (d > T(1) ? d : T(1));
        sm[ou[l] + k] = (sm[ou[l] + k] + e * m) * m;
      }
    __syncthreads();
    S.sweeps(L, ou[l], ot[l], ob[l], a.n_post, false, a.w[1], tx, ty);
  }
  // Level lc's corrected iterate back to global memory.
  {
    const LevelDev<T>& G = a.lv[a.lc];
    for (int j = ty; j < G.ny; j += 32)
      for (int i = tx; i < G.nx; i += 32) G.u[j * G.nx + i] = sm[ou[a.lc] + j * G.nx + i];
  }
}
template <typename T, int KIND>
__global__ __launch_bounds__(kCoarseThreads) void mg_coarse_lds_kernel(CoarseArgs<T> a) {
  extern __shared__ __align__(16) unsigned char coarse_smem[];
  mg_coarse_lds_body<T, KIND>(a, coarse_smem);
}

// Dynamic LDS of the coarse kernel for levels lc .. nl-1 (u, t, b each).
template <typename T>
size_t coarse_lds_bytes(const CoarseArgs<T>& a) {
  size_t n = 0;
  for (int l = a.lc; l < a.nl; ++l) n += 3 * (size_t)a.lv[l].ny * a.lv[l].nx;
  return n * sizeof(T);
}
// Dynamic LDS available to the coarse kernel: 160 KB less its static offset
// table (3 x 16 ints).
constexpr size_t kCoarseLdsMax = 160 * 1024 - 3 * 16 * sizeof(int);

template <typename T, int KIND>
static void launch_coarse(const CoarseArgs<T>& ca, hipStream_t s) {
  const size_t lds = coarse_lds_bytes(ca);
  if (lds <= kCoarseLdsMax)
    hipLaunchKernelGGL((mg_coarse_lds_kernel<T, KIND>), dim3(1), dim3(kCoarseThreads), lds, s,
                       ca);
  else
    hipLaunchKernelGGL((mg_coarse_kernel<T, KIND, false>), dim3(1), dim3(kCoarseThreads), 0, s,
                       ca);
}

// Allow the LDS-resident coarse kernels the full 160 KB (once per process).
static void coarse_lds_attributes() {
  static bool done = false;
  if (done) return;
  done = true;
  const int big = (int)kCoarseLdsMax;
  (void)hipFuncSetAttribute((const void*)mg_coarse_lds_kernel<float, kDirichlet>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, big);
  (void)hipFuncSetAttribute((const void*)mg_coarse_lds_kernel<float, kNoflux>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, big);
  (void)hipFuncSetAttribute((const void*)mg_coarse_lds_kernel<double, kDirichlet>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, big);
  (void)hipFuncSetAttribute((const void*)mg_coarse_lds_kernel<double, kNoflux>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, big);
}

// ---------------------------------------------------------------------------
// Pointwise / matvec kernels on one level (global memory, 64 x 4 blocks).
// ---------------------------------------------------------------------------
template <typename T, int KIND>
__global__ __launch_bounds__(256) void apply_op_kernel(LevelDev<T> L, const T* __restrict__ u,
                                                       T* __restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= L.nx || j >= L.ny) return;
  // The reference masks u before the stencil (u_m = u * mask).
  out[(int64_t)j * L.nx + i] = glob_Au<T, KIND>(L, u, j, i);
}

// res = (b - A u) * mask on one level.
template <typename T, int KIND>
__global__ __launch_bounds__(256) void residual_kernel(LevelDev<T> L, const T* __restrict__ u,
                                                       const T* __restrict__ b,
                                                       T* __restrict__ res) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= L.nx || j >= L.ny) return;
  const int64_t k = (int64_t)j * L.nx + i;
  res[k] = (b[k] - glob_Au<T, KIND>(L, u, j, i)) * mval<T>(L.mask, k);
}


template <typename T>
__global__ __launch_bounds__(256) void mask_copy_kernel(LevelDev<T> L, const T* __restrict__ in,
                                                        T* __restrict__ out, int64_t ld_in) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= L.nx || j >= L.ny) return;
  const int64_t k = (int64_t)j * L.nx + i;
  out[k] = in ? in[(int64_t)j * ld_in + i] * mval<T>(L.mask, k) : T(0);
}

template <typename T>
__global__ __launch_bounds__(256) void restrict_kernel(LevelDev<T> L, LevelDev<T> C,
                                                       const T* __restrict__ v,
                                                       T* __restrict__ out, int dirichlet) {
  const int I = blockIdx.x * 64 + threadIdx.x, J = blockIdx.y * 4 + threadIdx.y;
  if (I >= C.nx || J >= C.ny) return;
  T s4 = T(0), cnt = T(0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int dj = (q == 1 || q == 3), di = (q >= 2);
    const int64_t f = (int64_t)(2 * J + dj) * L.nx + (2 * I + di);
    const T m = mval<T>(L.mask, f);
    s4 += v[f] * m;
    cnt += m;
  }
  const int64_t g = (int64_t)J * C.nx + I;
  const T rs = dirichlet ? s4 * T(0.25) : s4 / (cnt > T(1) ? cnt : T(1));
  out[g] = rs * mval<T>(C.mask, g);
}

template <typename T, int KIND>
__global__ __launch_bounds__(256) void prolong_kernel(LevelDev<T> L, LevelDev<T> C,
                                                      T* __restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= L.nx || j >= L.ny) return;
  const int64_t k = (int64_t)j * L.nx + i;
  out[k] = prolong_at<T, KIND>(C, j, i) * mval<T>(L.mask, k);
}

// PCG: r = (b - A x) m, ||b||^2, ||r||^2; sets the stopping threshold.
template <typename T, int KIND>
__global__ __launch_bounds__(256) void pcg_init_kernel(LevelDev<T> L, const T* __restrict__ b,
                                                       int64_t ld_b, const T* __restrict__ x,
                                                       int64_t ld_x, T* __restrict__ r,
                                                       double rtol2, double atol2,
                                                       double* partials, PcgScalars* sc) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  double bb = 0.0, rr = 0.0;
  if (i < L.nx && j < L.ny) {
    const int64_t k = (int64_t)j * L.nx + i;
    const T m = mval<T>(L.mask, k);
    const T bv = b[(int64_t)j * ld_b + i] * m;
    const T ax = x ? glob_Au_pitched<T, KIND>(L, x, ld_x, j, i) : T(0);
    const T rv = (bv - ax) * m;
    r[k] = rv;
    bb = (double)bv * bv;
    rr = (double)rv * rv;
  }
  double v[2] = {bb, rr};
  block_partials<2, 256>(v, partials);
}

// PCG with a preconditioner other than the V-cycle: a = A z and the PCG dots
// (as the V-cycle's last kernel forms them).
// Each wave owns a 64-column strip of kStencilRows rows and slides a
// three-row register window down it (every row of z read once per strip);
// east/west neighbours come from lane shuffles, a strip's two edge columns
// from memory.  The four waves of a block sit side by side (256 columns), so
// three of their four edge loads hit the cache lines their neighbour wave on
// the same CU just read.
constexpr int kDirRows = 8;       // rows per wave of the pointwise update kernel
constexpr int kStencilRows = 8;   // rows per wave of the stencil kernel
static_assert(kDirRows == kStencilRows, "pcg_cg_update_az_kernel runs on grid_stencil");
static inline dim3 grid_dir(int ny, int nx) { return dim3(cdiv(nx, 64), cdiv(ny, 4 * kDirRows)); }
static inline dim3 grid_stencil(int ny, int nx) {
  return dim3(cdiv(nx, 256), cdiv(ny, kStencilRows));
}

template <typename T, int KIND>
__global__ __launch_bounds__(256) void pcg_az_dots_kernel(LevelDev<T> L, const T* __restrict__ z,
                                                          const T* __restrict__ r,
                                                          const T* __restrict__ q,
                                                          T* __restrict__ az, double* partials,
                                                          PcgScalars* sc) {
  const int done = sc->done;  // checked once the first rows are in flight
  const int lane = threadIdx.x;
  const int i = (blockIdx.x * 4 + threadIdx.y) * 64 + lane;
  const int j0 = blockIdx.y * kStencilRows;  // uniform within a wave
  const bool first = sc->first;
  const int nx = L.nx, ny = L.ny;
  const bool col = i < nx;
  // z at (row j, this lane's column); rows outside the level are zero
  // (j is uniform across the wave, col per lane).  32-bit offsets.
  auto Zrow = [&](int j) -> T { return (j < 0 || j >= ny || !col) ? T(0) : z[j * nx + i]; };
  auto Zat = [&](int j, int ii) -> T {  // edge lanes' outer neighbours
    return (ii < 0 || ii >= nx) ? T(0) : z[j * nx + ii];
  };
  auto CY = [&](int j) -> T {  // north-face coefficient of row j
    return (KIND == kNoflux && j >= 0 && j < ny && col) ? L.cy[j * nx + i] : T(0);
  };
  double d[kNumCgDots];
#pragma unroll
  for (int k = 0; k < kNumCgDots; ++k) d[k] = 0.0;
  T zs = Zrow(j0 - 1), zc = Zrow(j0);
  T cys = CY(j0 - 1);
  if (done) return;
  for (int rw = 0; rw < kStencilRows; ++rw) {
    const int j = j0 + rw;
    if (j >= ny) break;
    const T zn = Zrow(j + 1);
    T w = __shfl_up(zc, 1, 64);
    T e = __shfl_down(zc, 1, 64);
    if (lane == 0) w = Zat(j, i - 1);
    if (lane == 63) e = Zat(j, i + 1);
    T lap;
    T cyn = T(0);
    const int k = j * nx + i;
    if constexpr (KIND == kDirichlet) {
      lap = (e + w - T(2) * zc) * L.idx2 + (zn + zs - T(2) * zc) * L.idy2;
    } else {
      const T cxe = col ? L.cx[k] : T(0);
      T cxw = __shfl_up(cxe, 1, 64);
      if (lane == 0) cxw = (i > 0) ? L.cx[k - 1] : T(0);
      cyn = CY(j);
      lap = (cxe * (e - zc) - cxw * (zc - w)) * L.idx2 + (cyn * (zn - zc) - cys * (zc - zs)) * L.idy2;
    }
    if (col) {
      const T av = (lap - L.lam * zc) * mval<T>(L.mask, k);
      az[k] = av;
      const double zd = zc, rd = r[k], ad = av, qd = first ? 0.0 : (double)q[k];
      d[kZR] += zd * rd;
      d[kZQ] += zd * qd;
      d[kZA] += zd * ad;
      d[kRR] += rd * rd;
      d[kRA] += rd * ad;
      d[kRQ] += rd * qd;
      d[kAA] += ad * ad;
      d[kAQ] += ad * qd;
      d[kQQ] += qd * qd;
    }
    zs = zc;
    zc = zn;
    cys = cyn;
  }
  block_partials<kNumCgDots, 256>(d, partials);
}

// PCG step with the scalars of the finishing kernel (kRedCG):
//   p = z + beta p ; q = a + beta q ; x += alpha p ; r -= alpha q
// in place, pointwise.  Each wave updates a 64-column strip of kDirRows rows
// (grid_dir).  Runs unless done == 1 (done == 2: this is the last update).
template <typename T>
__device__ __forceinline__ void pcg_cg_update_body(int ny, int nx, T* __restrict__ x,
                                                            int64_t ld_x, const T* __restrict__ z,
                                                            const T* __restrict__ az,
                                                            T* __restrict__ p, T* __restrict__ q,
                                                            T* __restrict__ r, PcgScalars* sc, int bx, int by) {
  // The guard first: budgeted iterations after convergence then cost no
  // traffic (this kernel is bandwidth-bound, so the flag's round trip hides
  // behind the other blocks' loads).
  if (sc->done == 1) return;
  const int i = bx * 64 + threadIdx.x;
  const int j0 = (by * 4 + threadIdx.y) * kDirRows;
  const double beta64 = sc->beta;
  const T alpha = (T)sc->alpha, beta = (T)beta64;
  T xv[kDirRows], zv[kDirRows], av[kDirRows], pv[kDirRows], qv[kDirRows], rv[kDirRows];
#pragma unroll
  for (int t = 0; t < kDirRows; ++t) {
    const int j = j0 + t;
    const bool in = i < nx && j < ny;
    const int64_t k = (int64_t)j * nx + i;
    xv[t] = in ? x[(int64_t)j * ld_x + i] : T(0);
    zv[t] = in ? z[k] : T(0);
    av[t] = in ? az[k] : T(0);
    pv[t] = in ? p[k] : T(0);
    qv[t] = in ? q[k] : T(0);
    rv[t] = in ? r[k] : T(0);
  }
  // beta == 0 on the first step: p_{k-1}, q_{k-1} are not read into it.
  const bool restart = beta64 == 0.0;
#pragma unroll
  for (int t = 0; t < kDirRows; ++t) {
    const int j = j0 + t;
    if (i < nx && j < ny) {
      const int64_t k = (int64_t)j * nx + i;
      const T pn = restart ? zv[t] : zv[t] + beta * pv[t];
      const T qn = restart ? av[t] : av[t] + beta * qv[t];
      p[k] = pn;
      q[k] = qn;
      x[(int64_t)j * ld_x + i] = xv[t] + alpha * pn;
      r[k] = rv[t] - alpha * qn;
    }
  }
}
template <typename T>
__global__ __launch_bounds__(256) void pcg_cg_update_kernel(int ny, int nx, T* __restrict__ x,
                                                            int64_t ld_x, const T* __restrict__ z,
                                                            const T* __restrict__ az,
                                                            T* __restrict__ p, T* __restrict__ q,
                                                            T* __restrict__ r, PcgScalars* sc) {
  pcg_cg_update_body<T>(ny, nx, x, ld_x, z, az, p, q, r, sc, blockIdx.x, blockIdx.y);
}

// The same PCG step for FP32 levels without a land mask under the Dirichlet
// operator, whose wavefront-tile up leg does not store a = A z: a is formed
// here from z (three-row register window per 64-column strip, east/west
// neighbours through lane moves, the strip's edge columns from memory; the
// four waves of a block sit side by side as in grid_stencil, so most edge
// loads hit lines a neighbouring wave just read), which saves writing and
// reading one array per iteration.
template <typename T>
__global__ __launch_bounds__(256) void pcg_cg_update_az_kernel(
    LevelDev<T> L, T* __restrict__ x, int64_t ld_x, const T* __restrict__ z, T* __restrict__ p,
    const T* q, T* q_out, const T* r, T* r_out, PcgScalars* sc) {
  // q_out, r_out: q and r in place, or other arrays (each cell reads only
  // its own q and r).
  if (sc->done == 1) return;  // see pcg_cg_update_kernel
  const int lane = threadIdx.x;
  const int i = (blockIdx.x * 4 + threadIdx.y) * 64 + lane;
  const int j0 = blockIdx.y * kDirRows;  // grid_stencil with kDirRows == kStencilRows
  const int nx = L.nx, ny = L.ny;
  const bool col = i < nx;
  const double beta64 = sc->beta;
  const T alpha = (T)sc->alpha, beta = (T)beta64;
  auto Zat = [&](int j, int ii) -> T {
    return (j >= 0 && j < ny && ii >= 0 && ii < nx) ? z[(int64_t)j * nx + ii] : T(0);
  };
  T zr[kDirRows + 2], zw[kDirRows], ze[kDirRows];
  T xv[kDirRows], pv[kDirRows], qv[kDirRows], rv[kDirRows];
#pragma unroll
  for (int t = 0; t < kDirRows + 2; ++t) zr[t] = col ? Zat(j0 - 1 + t, i) : T(0);
#pragma unroll
  for (int t = 0; t < kDirRows; ++t) {
    const int j = j0 + t;
    const bool in = col && j < ny;
    const int64_t k = (int64_t)j * nx + i;
    xv[t] = in ? x[(int64_t)j * ld_x + i] : T(0);
    pv[t] = in ? p[k] : T(0);
    qv[t] = in ? q[k] : T(0);
    rv[t] = in ? r[k] : T(0);
    zw[t] = (lane == 0) ? Zat(j, i - 1) : T(0);
    ze[t] = (lane == 63) ? Zat(j, i + 1) : T(0);
  }
  const bool restart = beta64 == 0.0;
#pragma unroll
  for (int t = 0; t < kDirRows; ++t) {
    const int j = j0 + t;
    const T c = zr[t + 1];
    T w = lane_west(c), e = lane_east(c);
    if (lane == 0) w = zw[t];
    if (lane == 63) e = ze[t];
    const T lap = (e + w - T(2) * c) * L.idx2 + (zr[t + 2] + zr[t] - T(2) * c) * L.idy2;
    T av = lap - L.lam * c;  // z is zero on land (z m = z)
    if (col && j < ny) {
      const int64_t k = (int64_t)j * nx + i;
      if (L.mask && !L.mask[k]) av = T(0);  // (lap z - lam z) m
      const T pn = restart ? c : c + beta * pv[t];
      const T qn = restart ? av : av + beta * qv[t];
      p[k] = pn;
      q_out[k] = qn;
      x[(int64_t)j * ld_x + i] = xv[t] + alpha * pn;
      r_out[k] = rv[t] - alpha * qn;
    }
  }
}

__device__ __forceinline__ void scalars_clear(PcgScalars& x) {
  x.rz = x.pq = x.rr = x.bb = x.tol2 = x.alpha = x.beta = x.zq = x.out_rr = x.out_tol2 = 0.0;
  x.iter = x.done = x.first = x.out_done = x.outer = x.total = x.max_inner = 0;
  x.it_outer[0] = x.it_outer[1] = x.it_outer[2] = x.it_outer[3] = 0;
}

// Mixed precision: r32 = (b64 - A64 x64) m in FP64, ||r||^2 and ||b||^2 in
// FP64; the inner FP32 solve starts from e = 0 so e is cleared here too.
// Each wave owns a 64-column strip of kResRows rows and slides a three-row
// register window of x down it (every row of x read once per strip);
// east/west neighbours come from lane shuffles, a strip's two edge columns
// from memory.  Grid: grid_res (cdiv(nx, 64) x cdiv(ny, 4 kResRows)).
constexpr int kResRows = 8;
// Dirichlet operators use 62 output columns per wave (the two edge lanes only
// load the neighbouring columns).
static inline dim3 grid_res(int ny, int nx, int kind) {
  return dim3(cdiv(nx, kind == kDirichlet ? 62 : 64), cdiv(ny, 4 * kResRows));
}

template <int KIND, typename TB>
__device__ __forceinline__ void refine_residual_body(
    LevelDev<double> L, const TB* __restrict__ b, int64_t ld_b, const double* __restrict__ x,
    int64_t ld_x, float* __restrict__ r32, float* __restrict__ e32, double rtol2, double atol2,
    double* partials, PcgScalars* sc, int bx, int by, int gx, int gy, bool first = false) {
  // first: the residual that starts a new solve; one thread clears the
  // slot's inner and outer scalars (sc - 1, sc) and nobody reads them here.
  if (first && bx == 0 && by == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    scalars_clear(sc[-1]);
    scalars_clear(sc[0]);
  }
  const int lane = threadIdx.x;
  const int i = KIND == kDirichlet ? bx * 62 - 1 + lane : bx * 64 + lane;
  const int j0 = (by * 4 + threadIdx.y) * kResRows;  // uniform within a wave
  const int nx = L.nx, ny = L.ny;
  const bool col = i >= 0 && i < nx;
  // x at (row j, column ii), masked for the Dirichlet operator (the reference
  // applies the stencil to u * mask, multigrid.py:383); zero outside.
  auto Xat = [&](int j, int ii) -> double {
    if (j < 0 || j >= ny || ii < 0 || ii >= nx) return 0.0;
    const double v = x[(int64_t)j * ld_x + ii];
    return KIND == kDirichlet ? v * mval<double>(L.mask, (int64_t)j * nx + ii) : v;
  };
  auto CY = [&](int j) -> double {  // north-face coefficient of row j
    return (KIND == kNoflux && j >= 0 && j < ny && col) ? L.cy[(int64_t)j * nx + i] : 0.0;
  };
  if constexpr (KIND == kDirichlet) {
    // Lanes 1..62 own a column each, lanes 0 and 63 load the neighbouring
    // columns; every load of the strip is issued before any arithmetic (the
    // rows are independent), so each wave keeps ~30 loads in flight instead
    // of walking its rows one latency at a time.
    constexpr int R = kResRows;
    const bool own = col && lane >= 1 && lane <= 62;
    const int ic = min(max(i, 0), nx - 1);
    // bit t: the mask of row j0 - 1 + t in this lane's column, one byte load
    // per row for both the masked x and the residual's mask
    // (from two words of the column-packed mask; rows outside the level
    // read as dry, and their x and b are zero anyway)
    static_assert(R + 2 <= 32, "one 32-bit mask word per lane");
    uint32_t mb = ~0u;
    if (L.mcol) {
      const int g0 = j0 - 1, g = max(g0, 0);
      const int64_t blk = g >> 5;
      const uint64_t w0 = L.mcol[blk * nx + ic];
      const uint64_t w1 = L.mcol[(blk + 1) * nx + ic];
      mb = (uint32_t)((((w1 << 32) | w0) >> (g & 31)) << (g - g0));
    }
    double X[R + 2], B[R];
    float M[R];
#pragma unroll
    for (int t = 0; t < R + 2; ++t) {
      const int j = j0 - 1 + t;
      const double v = (col && j >= 0 && j < ny) ? x[(int64_t)j * ld_x + i] : 0.0;
      X[t] = ((mb >> t) & 1u) ? v : 0.0;  // x * mask
    }
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const int j = j0 + t;
      const bool in = own && j < ny;
      const int jc = min(j, ny - 1);
      M[t] = (in && ((mb >> (t + 1)) & 1u)) ? 1.0f : 0.0f;
      B[t] = in ? (double)b[(int64_t)jc * ld_b + ic] : 0.0;
    }
    // A certified solve has not moved x since the residual that certified it
    // (its updates are skipped), so the finishing kernel keeps that result.
    if (!first && sc->out_done) return;
    double bb = 0.0, rr = 0.0;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const int j = j0 + t;
      const double c = X[t + 1];
      const double w = dpp_move<0x138>(c);
      const double e = dpp_move<0x130>(c);
      const double lap = (e + w - 2.0 * c) * L.idx2 + (X[t + 2] + X[t] - 2.0 * c) * L.idy2;
      const double m = (double)M[t];
      const double bv = B[t] * m;
      const double rv = (bv - (lap - L.lam * c) * m) * m;
      if (own && j < ny) {
        const int64_t k = (int64_t)j * nx + i;
        r32[k] = (float)rv;
        e32[k] = 0.0f;
      }
      bb += bv * bv;
      rr += rv * rv;
    }
    double v[2] = {bb, rr};
    block_partials<2, 256>(v, partials, (int64_t)gx * gy, (int64_t)by * gx + bx);
    return;
  }
  const int done = first ? 0 : sc->out_done;  // checked once the first rows are in flight
  double xs = col ? Xat(j0 - 1, i) : 0.0, xc = col ? Xat(j0, i) : 0.0;
  double cys = CY(j0 - 1);
  // A certified solve has not moved x since the residual that certified it
  // (its updates are skipped), so the finishing kernel keeps that result.
  if (done) return;
  double bb = 0.0, rr = 0.0;
#pragma unroll
  for (int rw = 0; rw < kResRows; ++rw) {
    const int j = j0 + rw;
    if (j >= ny) break;  // wave-uniform
    const double xn = col ? Xat(j + 1, i) : 0.0;
    // lane - 1 and lane + 1 by DPP wave shifts (every lane active here)
    double w = dpp_move<0x138>(xc);
    double e = dpp_move<0x130>(xc);
    if (lane == 0) w = Xat(j, i - 1);
    if (lane == 63) e = Xat(j, i + 1);
    const int64_t k = (int64_t)j * nx + i;
    const double m = col ? mval<double>(L.mask, k) : 0.0;
    // centre u * mask; for the no-flux operator x itself is unmasked above
    const double c = (KIND == kDirichlet) ? xc : xc * m;
    double lap;
    double cyn = 0.0;
    if constexpr (KIND == kDirichlet) {
      lap = (e + w - 2.0 * c) * L.idx2 + (xn + xs - 2.0 * c) * L.idy2;
    } else {
      const double cxe = col ? L.cx[k] : 0.0;
      double cxw = dpp_move<0x138>(cxe);
      if (lane == 0) cxw = (i > 0 && i - 1 < nx) ? L.cx[k - 1] : 0.0;
      cyn = CY(j);
      lap = (cxe * (e - c) - cxw * (c - w)) * L.idx2 + (cyn * (xn - c) - cys * (c - xs)) * L.idy2;
    }
    if (col) {
      const double bv = (double)b[(int64_t)j * ld_b + i] * m;
      const double rv = (bv - (lap - L.lam * c) * m) * m;
      r32[k] = (float)rv;
      e32[k] = 0.0f;
      bb += bv * bv;
      rr += rv * rv;
    }
    xs = xc;
    xc = xn;
    cys = cyn;
  }
  double v[2] = {bb, rr};
  block_partials<2, 256>(v, partials, (int64_t)gx * gy, (int64_t)by * gx + bx);
}
template <int KIND, typename TB>
__global__ __launch_bounds__(256) void refine_residual_kernel(
    LevelDev<double> L, const TB* __restrict__ b, int64_t ld_b, const double* __restrict__ x,
    int64_t ld_x, float* __restrict__ r32, float* __restrict__ e32, double rtol2, double atol2,
    double* partials, PcgScalars* sc, int first) {
  refine_residual_body<KIND, TB>(L, b, ld_b, x, ld_x, r32, e32, rtol2, atol2, partials, sc,
                                 blockIdx.x, blockIdx.y, gridDim.x, gridDim.y, first != 0);
}

__device__ __forceinline__ void refine_update_body(int ny, int nx, double* __restrict__ x,
                                                   int64_t ld_x, const float* __restrict__ e,
                                                   PcgScalars* sc, int bx, int by) {
  // x += e over kDirRows rows per wave (grid_dir), loads first.
  if (sc->out_done) return;
  const int i = bx * 64 + threadIdx.x;
  const int j0 = (by * 4 + threadIdx.y) * kDirRows;
  double xv[kDirRows];
  float ev[kDirRows];
#pragma unroll
  for (int t = 0; t < kDirRows; ++t) {
    const int j = j0 + t;
    const bool in = i < nx && j < ny;
    xv[t] = in ? x[(int64_t)j * ld_x + i] : 0.0;
    ev[t] = in ? e[(int64_t)j * nx + i] : 0.0f;
  }
#pragma unroll
  for (int t = 0; t < kDirRows; ++t) {
    const int j = j0 + t;
    if (i < nx && j < ny) x[(int64_t)j * ld_x + i] = xv[t] + (double)ev[t];
  }
  if (i == 0 && j0 == 0) sc->outer += 1;
}
__global__ __launch_bounds__(256) void refine_update_kernel(int ny, int nx, double* __restrict__ x,
                                                            int64_t ld_x,
                                                            const float* __restrict__ e,
                                                            PcgScalars* sc) {
  refine_update_body(ny, nx, x, ld_x, e, sc, blockIdx.x, blockIdx.y);
}

// The same update when the inner solve's last PCG step is still pending (its
// V-cycles applied every other step inside their first kernel): the inner
// correction is e + alpha (z + beta p), formed in FP32 exactly as the step
// kernel forms x, then added in FP64 -- the step's r, q and p are not needed
// after an inner solve.  Skips the step when the inner solve stopped with it
// already applied (done == 1).
__device__ __forceinline__ void refine_update_step_body(
    int ny, int nx, double* __restrict__ x, int64_t ld_x, const float* __restrict__ e,
    const float* __restrict__ z, const float* __restrict__ p, const PcgScalars* in,
    PcgScalars* sc, int bx, int by) {
  if (sc->out_done) return;
  const bool step = in->done != 1;
  const double beta64 = in->beta;
  const float alpha = (float)in->alpha, beta = (float)beta64;
  const bool restart = beta64 == 0.0;
  const int i = bx * 64 + threadIdx.x;
  const int j0 = (by * 4 + threadIdx.y) * kDirRows;
  double xv[kDirRows];
  float ev[kDirRows], zv[kDirRows], pv[kDirRows];
#pragma unroll
  for (int t = 0; t < kDirRows; ++t) {
    const int j = j0 + t;
    const bool in_ = i < nx && j < ny;
    const int64_t k = (int64_t)j * nx + i;
    xv[t] = in_ ? x[(int64_t)j * ld_x + i] : 0.0;
    ev[t] = in_ ? e[k] : 0.0f;
    zv[t] = (in_ && step) ? z[k] : 0.0f;
    pv[t] = (in_ && step && !restart) ? p[k] : 0.0f;
  }
#pragma unroll
  for (int t = 0; t < kDirRows; ++t) {
    const int j = j0 + t;
    if (i < nx && j < ny) {
      float ef = ev[t];
      if (step) ef = ef + alpha * (restart ? zv[t] : zv[t] + beta * pv[t]);
      x[(int64_t)j * ld_x + i] = xv[t] + (double)ef;
    }
  }
  if (i == 0 && j0 == 0) sc->outer += 1;
}
__global__ __launch_bounds__(256) void refine_update_step_kernel(
    int ny, int nx, double* __restrict__ x, int64_t ld_x, const float* __restrict__ e,
    const float* __restrict__ z, const float* __restrict__ p, const PcgScalars* in,
    PcgScalars* sc) {
  refine_update_step_body(ny, nx, x, ld_x, e, z, p, in, sc, blockIdx.x, blockIdx.y);
}

// Reset an inner solve on an already-formed residual r (mixed precision):
// tol from ||r0||, flags cleared.  Single thread.
__global__ void inner_reset_kernel(PcgScalars* inner, const PcgScalars* outer, double inner_rtol) {
  inner_reset(inner, outer, inner_rtol);
}

// Certification: ||b - A x||^2 and ||b||^2 in FP64 into ``res`` (out_rr, bb,
// out_tol2, out_done).  Reads b and x pitched; x is not modified.
template <typename T, int KIND>
__global__ __launch_bounds__(256) void true_residual_kernel(LevelDev<T> L, const T* __restrict__ b,
                                                            int64_t ld_b, const T* __restrict__ x,
                                                            int64_t ld_x, double rtol2,
                                                            double atol2, double* partials,
                                                            PcgScalars* res) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  double bb = 0.0, rr = 0.0;
  if (i < L.nx && j < L.ny) {
    const int64_t k = (int64_t)j * L.nx + i;
    const double m = (double)mval<T>(L.mask, k);
    const double bv = (double)b[(int64_t)j * ld_b + i] * m;
    const double rv = (bv - (double)glob_Au_pitched<T, KIND>(L, x, ld_x, j, i)) * m;
    bb = bv * bv;
    rr = rv * rv;
  }
  double v[2] = {bb, rr};
  block_partials<2, 256>(v, partials);
}

// Clear the flags and counters of a scalar block (graph-safe replacement for
// a memset).
__global__ void scalars_reset_kernel(PcgScalars* sc, int n) {
  const int k = threadIdx.x;
  if (k < n) scalars_clear(sc[k]);
}

// ---------------------------------------------------------------------------
// Host-side plan
// ---------------------------------------------------------------------------

template <typename T>
struct Hier {
  std::vector<LevelDev<T>> lv;
  std::vector<void*> allocs;
  // PCG vectors on level 0
  // r residual, z = M r, az = A z, p direction, q = A p, x iterate
  T *r = nullptr, *z = nullptr, *p = nullptr, *az = nullptr, *q = nullptr, *x = nullptr;
  // second r and q (FP32): the PCG step fused into the next V-cycle's down
  // leg reads r_k, q_{k-1} on its tiles' halos and writes r_{k+1}, q_k here
  T *r2 = nullptr, *q2 = nullptr;
};

struct Plan {
  int kind, prec, nl, ny, nx;
  int n_pre, n_post, n_coarse;
  double omega, lam, dx, dy;
  // Per-sweep weights of the wave-tile sweeps: [0] pre-smoothing (down
  // legs), [1] post-smoothing (up legs), [2] the bottom level; omega unless
  // Chebyshev smoothing is on (plan_set_chebyshev).
  double omw[3][kMaxW];
  bool cheb = false;
  int lc;  // first level handled by the single-workgroup coarse kernel
  bool has64 = false, has32 = false;
  Hier<double> h64;
  Hier<float> h32;
  double* partials = nullptr;
  size_t npart = 0;  // capacity of partials in doubles
  PcgScalars* sc = nullptr;   // kSlots pairs: [2k] inner, [2k+1] outer
  PcgScalars* hsc = nullptr;  // pinned host mirror
  int last_iters = 4, last_outer = 2, last_inner = 4;
  float* work0[4] = {nullptr, nullptr, nullptr, nullptr};  // fine-level FP32 work (lazy)
  double* dscal = nullptr;  // device scalars of the learned correction (lazy)
  std::vector<void*> allocs;
};

static int compute_n_levels(int ny, int nx, int min_coarse = 8) {  // multigrid.py:50-76
  int n = 1;
  while (ny % 2 == 0 && nx % 2 == 0 && ny / 2 >= min_coarse && nx / 2 >= min_coarse) {
    ny /= 2;
    nx /= 2;
    ++n;
  }
  return n;
}

struct HostLevel {
  int ny, nx;
  double dx, dy;
  std::vector<double> mask, coeff, cx, cy, inv;
  bool all_wet;
};

static std::vector<HostLevel> build_host_levels(const double* mask, const double* coeff, int ny,
                                                int nx, double dx, double dy, double lam,
                                                int nl, int kind) {
  std::vector<HostLevel> out;
  HostLevel cur;
  cur.ny = ny;
  cur.nx = nx;
  cur.dx = dx;
  cur.dy = dy;
  const size_t n0 = (size_t)ny * nx;
  cur.mask.assign(mask, mask + n0);
  cur.coeff.resize(n0);
  for (size_t k = 0; k < n0; ++k) cur.coeff[k] = (coeff ? coeff[k] : 1.0) * cur.mask[k];
  for (int l = 0; l < nl; ++l) {
    const int Ny = cur.ny, Nx = cur.nx;
    const size_t N = (size_t)Ny * Nx;
    cur.all_wet = true;
    for (size_t k = 0; k < N; ++k)
      if (cur.mask[k] != 1.0) cur.all_wet = false;
    if (kind == kNoflux) {
      // _interpolate_coeff_to_faces + _compute_inv_diagonal (multigrid.py:143-279)
      cur.cx.assign(N, 0.0);
      cur.cy.assign(N, 0.0);
      for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
          const size_t k = (size_t)j * Nx + i;
          if (i + 1 < Nx)
            cur.cx[k] = 0.5 * (cur.coeff[k] + cur.coeff[k + 1]) * (cur.mask[k] * cur.mask[k + 1]);
          if (j + 1 < Ny)
            cur.cy[k] = 0.5 * (cur.coeff[k] + cur.coeff[k + Nx]) * (cur.mask[k] * cur.mask[k + Nx]);
        }
      cur.inv.assign(N, 0.0);
      for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
          const size_t k = (size_t)j * Nx + i;
          const double cxw = i > 0 ? cur.cx[k - 1] : 0.0;
          const double cys = j > 0 ? cur.cy[k - Nx] : 0.0;
          double d = -(cur.cx[k] + cxw) / (cur.dx * cur.dx) - (cur.cy[k] + cys) / (cur.dy * cur.dy) - lam;
          d *= cur.mask[k];
          cur.inv[k] = (std::fabs(d) > 1e-30 ? 1.0 / d : 0.0) * cur.mask[k];
        }
    }
    HostLevel next;
    if (l < nl - 1) {
      // _restrict_mask / _restrict_coeff (multigrid.py:79-101, 182-218)
      next.ny = Ny / 2;
      next.nx = Nx / 2;
      next.dx = cur.dx * 2.0;
      next.dy = cur.dy * 2.0;
      const size_t Nc = (size_t)next.ny * next.nx;
      next.mask.assign(Nc, 0.0);
      next.coeff.assign(Nc, 0.0);
      for (int J = 0; J < next.ny; ++J)
        for (int I = 0; I < next.nx; ++I) {
          const size_t a = (size_t)(2 * J) * Nx + 2 * I, b = a + Nx, c = a + 1, d = a + Nx + 1;
          const double avg = (cur.mask[a] + cur.mask[b] + cur.mask[c] + cur.mask[d]) / 4.0;
          const double mc = avg >= 0.5 ? 1.0 : 0.0;
          const double s4 = cur.coeff[a] * cur.mask[a] + cur.coeff[b] * cur.mask[b] +
                            cur.coeff[c] * cur.mask[c] + cur.coeff[d] * cur.mask[d];
          const double cnt = cur.mask[a] + cur.mask[b] + cur.mask[c] + cur.mask[d];
          next.mask[(size_t)J * next.nx + I] = mc;
          next.coeff[(size_t)J * next.nx + I] = (s4 / std::max(cnt, 1.0)) * mc;
        }
    }
    out.push_back(std::move(cur));
    if (l < nl - 1) cur = std::move(next);
  }
  return out;
}

template <typename T>
static int upload_hier(Plan& P, Hier<T>& H, const std::vector<HostLevel>& hl) {
  auto alloc = [&](size_t bytes, void** p) -> int {
    FVX_HIP(hipMalloc(p, bytes));
    H.allocs.push_back(*p);
    return kOk;
  };
  H.lv.resize(hl.size());
  for (size_t l = 0; l < hl.size(); ++l) {
    const HostLevel& h = hl[l];
    LevelDev<T>& L = H.lv[l];
    std::memset(&L, 0, sizeof(L));
    L.ny = h.ny;
    L.nx = h.nx;
    L.idx2 = (T)(1.0 / (h.dx * h.dx));
    L.idy2 = (T)(1.0 / (h.dy * h.dy));
    L.lam = (T)P.lam;
    L.inv_d = (T)(1.0 / (-2.0 / (h.dx * h.dx) - 2.0 / (h.dy * h.dy) - P.lam));
    const size_t N = (size_t)h.ny * h.nx;
    if (!h.all_wet) {
      std::vector<uint8_t> m(N);
      for (size_t k = 0; k < N; ++k) m[k] = h.mask[k] != 0.0;
      void* d;
      int rc = alloc(N, &d);
      if (rc) return rc;
      FVX_HIP(hipMemcpy(d, m.data(), N, hipMemcpyHostToDevice));
      L.mask = (const uint8_t*)d;
      const size_t nb = (size_t)h.ny / 32 + 2;
      std::vector<uint32_t> cb(nb * h.nx, 0u);
      for (int j = 0; j < h.ny; ++j)
        for (int i = 0; i < h.nx; ++i)
          if (m[(size_t)j * h.nx + i]) cb[(size_t)(j >> 5) * h.nx + i] |= 1u << (j & 31);
      rc = alloc(cb.size() * sizeof(uint32_t), &d);
      if (rc) return rc;
      FVX_HIP(hipMemcpy(d, cb.data(), cb.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      L.mcol = (const uint32_t*)d;
    }
    if (P.kind == kNoflux) {
      std::vector<T> tmp(N);
      const std::vector<double>* srcs[3] = {&h.cx, &h.cy, &h.inv};
      const T** dsts[3] = {&L.cx, &L.cy, &L.inv};
      for (int a = 0; a < 3; ++a) {
        for (size_t k = 0; k < N; ++k) tmp[k] = (T)(*srcs[a])[k];
        void* d;
        int rc = alloc(N * sizeof(T), &d);
        if (rc) return rc;
        FVX_HIP(hipMemcpy(d, tmp.data(), N * sizeof(T), hipMemcpyHostToDevice));
        *dsts[a] = (const T*)d;
      }
    }
    void* w[5];
    for (int a = 0; a < 5; ++a) {
      int rc = alloc(N * sizeof(T), &w[a]);
      if (rc) return rc;
      FVX_HIP(hipMemset(w[a], 0, N * sizeof(T)));
    }
    L.u = (T*)w[0];
    L.b = (T*)w[1];
    L.t = (T*)w[2];
    L.t2 = (T*)w[3];
    L.t3 = (T*)w[4];
  }
  const size_t N0 = (size_t)hl[0].ny * hl[0].nx;
  void* v[8] = {};
  const int nvec = std::is_same<T, float>::value ? 8 : 6;
  for (int k = 0; k < nvec; ++k) {
    int rc = alloc(N0 * sizeof(T), &v[k]);
    if (rc) return rc;
    FVX_HIP(hipMemset(v[k], 0, N0 * sizeof(T)));
  }
  H.r = (T*)v[0];
  H.z = (T*)v[1];
  H.p = (T*)v[2];
  H.az = (T*)v[3];
  H.q = (T*)v[4];
  H.x = (T*)v[5];
  H.r2 = (T*)v[6];
  H.q2 = (T*)v[7];
  return kOk;
}

template <typename T>
static void free_hier(Hier<T>& H) {
  for (void* p : H.allocs) (void)hipFree(p);
  H.allocs.clear();
  H.lv.clear();
}

// Cells at or below which a level is handled by the coarse single-workgroup kernel.
constexpr int kCoarseCells = 64 * 64;
// Independent solve slots (device scalar blocks) per plan.
constexpr int kSlots = 8;

template <typename T>
static Hier<T>& hier(Plan& P);
template <>
Hier<double>& hier<double>(Plan& P) { return P.h64; }
template <>
Hier<float>& hier<float>(Plan& P) { return P.h32; }

// Levels the wavefront-tile kernels run: FP32, Dirichlet operator, no land
// mask (nor on the coarser level: the restriction of an all-wet mask is all
// wet).  Their up leg with the PCG dots does not store a = A z; the PCG
// update kernel of such a level recomputes it (pcg_cg_update_az_kernel).
template <typename T, int KIND>
static bool wave_tiled(const LevelDev<T>& L, const LevelDev<T>* C) {
  // (its buffer accesses address each array in 32-bit byte offsets)
  (void)C;
  if constexpr (std::is_same<T, float>::value && KIND == kDirichlet)
    return (int64_t)L.ny * L.nx < (int64_t(1) << 29);
  return false;
}

template <typename T, int KIND, int NS, int MODE, bool ZERO>
static void launch_tile(const Plan& P, const LevelDev<T>& L, const LevelDev<T>* C, const T* u_in,
                        const T* b, T* u_out, const T* w, PcgScalars* sc, hipStream_t s,
                        T* az_out = nullptr, bool finish = true,
                        const PcgUpdArgs<T>* upd = nullptr, bool uniform = false) {
  TileArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  a.L = L;
  if (C) a.C = *C;
  a.u_in = u_in;
  a.b = b;
  a.u_out = u_out;
  a.w = w;
  a.az_out = az_out;
  a.omega = (T)P.omega;
  {
    // (sweep weights by leg: down, up, bottom; ``uniform``: omega throughout)
    constexpr int role = (MODE == kDown || MODE == kDownUpd) ? 0 : (MODE == kSmooth ? 2 : 1);
    for (int k = 0; k < 3; ++k) a.omw[k] = uniform ? a.omega : (T)P.omw[role][k];
  }
  a.partials = P.partials;
  a.sc = sc;
  if constexpr (std::is_same<T, float>::value && KIND == kDirichlet) {
    // (a plain smoothing call on a masked level keeps the iterate's land
    // values, as the reference does: the LDS kernel)
    if (wave_tiled<T, KIND>(L, C) && !(MODE == kSmooth && !ZERO && L.mask)) {
      // FP32 Dirichlet level: one wavefront per tile (mg_wave_tile_kernel)
      a.az_out = nullptr;  // recomputed by pcg_cg_update_az_kernel
      if (upd) a.upd = *upd;
      constexpr int HL = wt_lane_halo<NS, MODE>();
      constexpr bool kMayShort = MODE == kSmooth || MODE == kDown || MODE == kUp;
      const bool short_tiles = kMayShort && (int64_t)L.ny * L.nx <= kShortTileMaxCells;
      a.tiles_x = cdiv(L.nx, 64 - 2 * HL);
      constexpr int kLongRows = kMayShort ? kTileRows : wt_pcg_rows<MODE>();
      a.ntiles = a.tiles_x * cdiv(L.ny, short_tiles ? kShortTileRows : kLongRows);
      const int nblk = cdiv(a.ntiles, 4);
      const bool mk = L.mask || (C && C->mask);
      // the wall correction exists on coarse levels only (no PCG modes)
      constexpr bool kMayWall = MODE == kSmooth || MODE == kDown || MODE == kUp;
      const bool wall = kMayWall && L.wall_g != T(0);
      auto go = [&](auto rows, auto masked) {
        if constexpr (kMayWall) {
          if (wall) {
            hipLaunchKernelGGL((mg_wave_tile_kernel<NS, MODE, ZERO, decltype(rows)::value,
                                                    decltype(masked)::value, true>),
                               dim3(nblk), dim3(256), 0, s, a);
            return;
          }
        }
        hipLaunchKernelGGL((mg_wave_tile_kernel<NS, MODE, ZERO, decltype(rows)::value,
                                                decltype(masked)::value>),
                           dim3(nblk), dim3(256), 0, s, a);
      };
      using Masked = std::true_type;
      using Wet = std::false_type;
      using LongRows = std::integral_constant<int, kLongRows>;
      if constexpr (kMayShort) {
        using ShortRows = std::integral_constant<int, kShortTileRows>;
        if (short_tiles) mk ? go(ShortRows(), Masked()) : go(ShortRows(), Wet());
        else mk ? go(LongRows(), Masked()) : go(LongRows(), Wet());
      } else {
        mk ? go(LongRows(), Masked()) : go(LongRows(), Wet());
      }
      if constexpr (MODE == kUpCG)
        if (finish)
          hipLaunchKernelGGL((reduce_finish_kernel<kNumCgDots, kRedCG>), dim3(1),
                             dim3(kFinishThreads), 0, s, P.partials, nblk, sc, 0.0, 0.0);
      return;
    }
  }
  if constexpr (MODE != kDownUpd) {
    a.tiles_x = cdiv(L.nx, kTx);
    a.ntiles = a.tiles_x * cdiv(L.ny, kTy);
    if (L.mask)
      hipLaunchKernelGGL((mg_tile_kernel<T, KIND, NS, MODE, ZERO, true>), dim3(a.ntiles),
                         dim3(kTileThreads), 0, s, a);
    else
      hipLaunchKernelGGL((mg_tile_kernel<T, KIND, NS, MODE, ZERO, false>), dim3(a.ntiles),
                         dim3(kTileThreads), 0, s, a);
    if constexpr (MODE == kUpCG)
      if (finish)
        hipLaunchKernelGGL((reduce_finish_kernel<kNumCgDots, kRedCG>), dim3(1), dim3(kFinishThreads),
                           0, s, P.partials, a.ntiles, sc, 0.0, 0.0);
  }
}

// Damped-Jacobi sweeps fused into one level kernel (with the residual and
// restriction on the down leg, the prolongation on the up leg); more sweeps
// per side run as separate smoothing kernels first (down) or after (up).
constexpr int kMaxFuse = 3;

// Level-0 up kernel of a preconditioning V-cycle with the PCG dots, with the
// plan's fused post-sweeps, as the V-cycle launches it, without its
// finishing kernel (profiling hook).
template <typename T, int KIND>
static void launch_vcycle_up_cg(const Plan& P, const LevelDev<T>& L, const LevelDev<T>& C,
                                const T* u, const T* b, T* out, const T* w, T* az,
                                PcgScalars* sc, hipStream_t s) {
  const int fz = P.n_post < kMaxFuse ? P.n_post : kMaxFuse;
  if (fz == 3) launch_tile<T, KIND, 3, kUpCG, false>(P, L, &C, u, b, out, w, sc, s, az, false);
  else if (fz == 2) launch_tile<T, KIND, 2, kUpCG, false>(P, L, &C, u, b, out, w, sc, s, az, false);
  else if (fz == 1) launch_tile<T, KIND, 1, kUpCG, false>(P, L, &C, u, b, out, w, sc, s, az, false);
  else launch_tile<T, KIND, 0, kUpCG, false>(P, L, &C, u, b, out, w, sc, s, az, false);
}

// Finishing kernel of a reduction over the blocks of grid2(ny, nx).
template <int NV, int SITE>
static void finish2(const Plan& P, int ny, int nx, PcgScalars* sc, hipStream_t s,
                    double rtol2 = 0.0, double atol2 = 0.0) {
  const int nblk = cdiv(nx, 64) * cdiv(ny, 4);
  hipLaunchKernelGGL((reduce_finish_kernel<NV, SITE>), dim3(1), dim3(kFinishThreads), 0, s,
                     P.partials, nblk, sc, rtol2, atol2);
}

// nu Jacobi sweeps on level L, up to 3 per kernel.  Ping-pongs between the
// scratch arrays s1 / s2 (which must differ from u_in and final_out); the last
// chunk writes final_out when given.  Returns the array holding the result.
// The sweeps take the plan's bottom-level weights unless ``uniform`` (omega).
template <typename T, int KIND>
static T* smooth_chunks(const Plan& P, const LevelDev<T>& L, const T* u_in, bool zero,
                        const T* b, int nu, T* s1, T* s2, T* final_out, PcgScalars* sc,
                        hipStream_t s, bool uniform = false) {
  const T* src = u_in;
  T* dst = nullptr;
  const int K = (nu + 2) / 3;
  for (int c = 0; c < K; ++c) {
    const int ns = std::min(3, nu - 3 * c);
    dst = (c == K - 1 && final_out) ? final_out : ((c & 1) ? s2 : s1);
    const bool z = zero && c == 0;
    if (ns == 3) {
      if (z)
        launch_tile<T, KIND, 3, kSmooth, true>(P, L, nullptr, src, b, dst, nullptr, sc, s, nullptr,
                                                 true, nullptr, uniform);
      else
        launch_tile<T, KIND, 3, kSmooth, false>(P, L, nullptr, src, b, dst, nullptr, sc, s,
                                                  nullptr, true, nullptr, uniform);
    } else if (ns == 2) {
      if (z)
        launch_tile<T, KIND, 2, kSmooth, true>(P, L, nullptr, src, b, dst, nullptr, sc, s, nullptr,
                                                 true, nullptr, uniform);
      else
        launch_tile<T, KIND, 2, kSmooth, false>(P, L, nullptr, src, b, dst, nullptr, sc, s,
                                                  nullptr, true, nullptr, uniform);
    } else {
      if (z)
        launch_tile<T, KIND, 1, kSmooth, true>(P, L, nullptr, src, b, dst, nullptr, sc, s, nullptr,
                                                 true, nullptr, uniform);
      else
        launch_tile<T, KIND, 1, kSmooth, false>(P, L, nullptr, src, b, dst, nullptr, sc, s,
                                                  nullptr, true, nullptr, uniform);
    }
    src = dst;
  }
  return dst;
}

// One V-cycle z = V(u0 or 0, rhs) from level l0 (rhs, u0 and out have that
// level's shape; MultigridSolver.v_cycle(u, rhs, level_idx), multigrid.py:
// 808-883).  With ``az`` given (level 0 only), the last kernel also writes
// az = A z and the PCG dots against rhs and w (kUpCG) when it can (two or
// fewer post-sweeps on a tiled level 0); *fused says whether it did.
// The coarse kernel's sweep weights (omega, or the plan's Chebyshev weights).
template <typename T>
static void coarse_weights(const Plan& P, CoarseArgs<T>& ca) {
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < kMaxW; ++k) ca.w[r][k] = (T)P.omw[r][k];
}

template <typename T, int KIND>
static int v_cycle_impl(Plan& P, const T* u0, const T* rhs, T* out, T* az, const T* w,
                        PcgScalars* sc, hipStream_t s, bool* fused,
                        const PcgUpdArgs<T>* upd = nullptr, int l0 = 0) {
  Hier<T>& H = hier<T>(P);
  const int nl = P.nl, lc = P.lc;
  if (fused) *fused = false;
  if (l0 >= lc) {
    const LevelDev<T>& L = H.lv[l0];
    const size_t bytes = sizeof(T) * (size_t)L.ny * L.nx;
    if ((int64_t)L.ny * L.nx > kCoarseCells) {
      // One large level (grid not divisible): the V-cycle is n_coarse sweeps.
      if (P.n_coarse == 0) {
        if (u0) FVX_HIP(hipMemcpyAsync(out, u0, bytes, hipMemcpyDeviceToDevice, s));
        else FVX_HIP(hipMemsetAsync(out, 0, bytes, s));
      } else {
        smooth_chunks<T, KIND>(P, L, u0, u0 == nullptr, rhs, P.n_coarse, L.t, L.t2, out, sc, s);
      }
      FVX_LAUNCH_CHECK();
      return kOk;
    }
    // Levels l0 .. nl-1 in the single-workgroup kernel.
    FVX_HIP(hipMemcpyAsync(L.b, rhs, bytes, hipMemcpyDeviceToDevice, s));
    if (u0) FVX_HIP(hipMemcpyAsync(L.u, u0, bytes, hipMemcpyDeviceToDevice, s));
    CoarseArgs<T> ca;
    std::memset(&ca, 0, sizeof(ca));
    for (int l = 0; l < nl; ++l) ca.lv[l] = H.lv[l];
    ca.lc = l0;
    ca.nl = nl;
    ca.n_pre = P.n_pre;
    ca.n_post = P.n_post;
    ca.n_coarse = P.n_coarse;
    ca.nonzero_init = u0 != nullptr;
    ca.omega = (T)P.omega;
    coarse_weights(P, ca);
    ca.done = sc ? &sc->done : nullptr;
    launch_coarse<T, KIND>(ca, s);
    FVX_HIP(hipMemcpyAsync(out, L.u, bytes, hipMemcpyDeviceToDevice, s));
    FVX_LAUNCH_CHECK();
    return kOk;
  }
  // Down leg over the tiled levels l0 .. lc-1: result iterate in L.u.
  for (int l = l0; l < lc; ++l) {
    const LevelDev<T>& L = H.lv[l];
    const LevelDev<T>& C = H.lv[l + 1];
    const T* b = (l == l0) ? rhs : L.b;
    bool zero = (l > l0) || (u0 == nullptr);
    const T* src = (l == l0) ? u0 : nullptr;
    const int npre = P.n_pre;
    const int fz = npre < kMaxFuse ? npre : kMaxFuse;  // fused with residual + restriction
    if (l == 0 && upd) {
      // The previous PCG step, then this down leg from r_{k+1} = rhs
      // (pcg_iters checks that the plan allows it).
      if (fz == 3) launch_tile<T, KIND, 3, kDownUpd, true>(P, L, &C, nullptr, upd->r_in, L.u,
                                                           nullptr, sc, s, nullptr, true, upd);
      else if (fz == 2) launch_tile<T, KIND, 2, kDownUpd, true>(P, L, &C, nullptr, upd->r_in, L.u,
                                                                nullptr, sc, s, nullptr, true, upd);
      else if (fz == 1) launch_tile<T, KIND, 1, kDownUpd, true>(P, L, &C, nullptr, upd->r_in, L.u,
                                                                nullptr, sc, s, nullptr, true, upd);
      else launch_tile<T, KIND, 0, kDownUpd, true>(P, L, &C, nullptr, upd->r_in, L.u, nullptr, sc,
                                                   s, nullptr, true, upd);
      continue;
    }
    if (npre - fz > 0) {
      src = smooth_chunks<T, KIND>(P, L, src, zero, b, npre - fz, L.t, L.t2, nullptr, sc, s);
      zero = false;
    }
    if (fz == 3) {
      if (zero) launch_tile<T, KIND, 3, kDown, true>(P, L, &C, src, b, L.u, nullptr, sc, s);
      else launch_tile<T, KIND, 3, kDown, false>(P, L, &C, src, b, L.u, nullptr, sc, s);
    } else if (fz == 2) {
      if (zero) launch_tile<T, KIND, 2, kDown, true>(P, L, &C, src, b, L.u, nullptr, sc, s);
      else launch_tile<T, KIND, 2, kDown, false>(P, L, &C, src, b, L.u, nullptr, sc, s);
    } else if (fz == 1) {
      if (zero) launch_tile<T, KIND, 1, kDown, true>(P, L, &C, src, b, L.u, nullptr, sc, s);
      else launch_tile<T, KIND, 1, kDown, false>(P, L, &C, src, b, L.u, nullptr, sc, s);
    } else {
      if (zero) launch_tile<T, KIND, 0, kDown, true>(P, L, &C, src, b, L.u, nullptr, sc, s);
      else launch_tile<T, KIND, 0, kDown, false>(P, L, &C, src, b, L.u, nullptr, sc, s);
    }
  }
  // Coarse levels lc .. nl-1 in one workgroup (zero guess at level lc); a
  // bottom level too large for it gets n_coarse damped-Jacobi sweeps from
  // zero on the tiled kernels instead.
  if (lc == nl - 1 && (int64_t)H.lv[lc].ny * H.lv[lc].nx > kCoarseCells) {
    const LevelDev<T>& B = H.lv[lc];
    smooth_chunks<T, KIND>(P, B, nullptr, true, B.b, std::max(P.n_coarse, 1), B.t, B.t2, B.u, sc,
                           s);
  } else if (lc < nl) {
    CoarseArgs<T> ca;
    std::memset(&ca, 0, sizeof(ca));
    for (int l = 0; l < nl; ++l) ca.lv[l] = H.lv[l];
    ca.lc = lc;
    ca.nl = nl;
    ca.n_pre = P.n_pre;
    ca.n_post = P.n_post;
    ca.n_coarse = P.n_coarse;
    ca.omega = (T)P.omega;
    coarse_weights(P, ca);
    ca.done = sc ? &sc->done : nullptr;
    launch_coarse<T, KIND>(ca, s);
  }
  // Up leg over the tiled levels lc-1 .. l0.  cur = array holding level l+1's
  // corrected iterate; each up kernel reads L.u and writes a different array.
  T* cur = (lc < nl) ? H.lv[lc].u : nullptr;
  for (int l = lc - 1; l >= l0; --l) {
    const LevelDev<T>& L = H.lv[l];
    LevelDev<T> C = H.lv[l + 1];
    C.u = cur;
    const T* b = (l == l0) ? rhs : L.b;
    const int npost = P.n_post;
    const int fz = npost < kMaxFuse ? npost : kMaxFuse;
    const int tail = npost - fz;
    const bool dk = az && l == 0 && tail == 0;
    T* dst = (l == l0 && tail == 0) ? out : L.t;
    if (dk) {
      if (fz == 3) launch_tile<T, KIND, 3, kUpCG, false>(P, L, &C, L.u, b, dst, w, sc, s, az);
      else if (fz == 2) launch_tile<T, KIND, 2, kUpCG, false>(P, L, &C, L.u, b, dst, w, sc, s, az);
      else if (fz == 1) launch_tile<T, KIND, 1, kUpCG, false>(P, L, &C, L.u, b, dst, w, sc, s, az);
      else launch_tile<T, KIND, 0, kUpCG, false>(P, L, &C, L.u, b, dst, w, sc, s, az);
      if (fused) *fused = true;
    } else {
      if (fz == 3) launch_tile<T, KIND, 3, kUp, false>(P, L, &C, L.u, b, dst, nullptr, sc, s);
      else if (fz == 2) launch_tile<T, KIND, 2, kUp, false>(P, L, &C, L.u, b, dst, nullptr, sc, s);
      else if (fz == 1) launch_tile<T, KIND, 1, kUp, false>(P, L, &C, L.u, b, dst, nullptr, sc, s);
      else launch_tile<T, KIND, 0, kUp, false>(P, L, &C, L.u, b, dst, nullptr, sc, s);
    }
    if (tail > 0) {
      dst = smooth_chunks<T, KIND>(P, L, dst, false, b, tail, L.t2, L.t3, l == l0 ? out : nullptr,
                                   sc, s);
    }
    cur = dst;
  }
  FVX_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
static int v_cycle(Plan& P, const T* u0, const T* rhs, T* out, PcgScalars* sc, hipStream_t s,
                   int l0 = 0) {
  return P.kind == kDirichlet
             ? v_cycle_impl<T, kDirichlet>(P, u0, rhs, out, nullptr, nullptr, sc, s, nullptr,
                                           nullptr, l0)
             : v_cycle_impl<T, kNoflux>(P, u0, rhs, out, nullptr, nullptr, sc, s, nullptr,
                                        nullptr, l0);
}

static dim3 grid2(int ny, int nx) { return dim3(cdiv(nx, 64), cdiv(ny, 4)); }

// Enqueue ``n`` flexible-PCG iterations (device-guarded) on hierarchy H.
// Precondition: H.r holds the residual and sc->first / tol2 are set.  One
// iteration is z = M r with az = A z and the PCG dots (fused into the
// V-cycle's last kernel, else a stencil kernel), the finishing kernel that
// turns the dots into the step's scalars, and the pointwise update.
template <typename T, int KIND>
static int pcg_iters(Plan& P, Hier<T>& H, T* x, int64_t ld_x, int n, PcgScalars* sc,
                     hipStream_t s, PrecondFn pf = nullptr, void* ctx = nullptr,
                     bool* defer_last = nullptr) {
  const LevelDev<T>& L = H.lv[0];
  if (defer_last) *defer_last = false;
  // All-wet FP32 level 0 under the V-cycle with fused sweeps: each PCG step
  // after the first is applied by the next V-cycle's level-0 down kernel
  // (kDownUpd), which reads r_k, q_{k-1} from one pair of arrays and writes
  // r_{k+1}, q_k to the other; the last step runs on its own and leaves r, q
  // in H.r, H.q.
  bool fuse_step = false;
  if constexpr (std::is_same<T, float>::value && KIND == kDirichlet)
    fuse_step = !pf && P.lc >= 1 && P.nl > 1 && wave_tiled<T, KIND>(L, &H.lv[1]) &&
                P.n_pre <= kMaxFuse && P.n_post <= kMaxFuse && H.r2 && H.q2;
  if (fuse_step) {
    if constexpr (std::is_same<T, float>::value && KIND == kDirichlet) {
      T *rc = H.r, *qc = H.q, *rn = H.r2, *qn = H.q2;
      for (int k = 0; k < n; ++k) {
        bool fused = false;
        int rc_;
        if (k == 0) {
          rc_ = v_cycle_impl<T, KIND>(P, nullptr, rc, H.z, H.az, qc, sc, s, &fused);
        } else {
          PcgUpdArgs<T> u{H.z, H.p, x, ld_x, qc, qn, rc, rn};
          rc_ = v_cycle_impl<T, KIND>(P, nullptr, rn, H.z, H.az, qn, sc, s, &fused, &u);
          std::swap(rc, rn);
          std::swap(qc, qn);
        }
        if (rc_) return rc_;
        if (!fused) {
          set_error("internal: the fused PCG step needs the V-cycle's up leg with the dots");
          return kInvalidArgument;
        }
      }
      // The last step: here, or (defer_last) by the caller, which needs only
      // its x (refine_update_step_kernel).
      if (defer_last) *defer_last = true;
      else
        hipLaunchKernelGGL(pcg_cg_update_az_kernel<T>, grid_stencil(L.ny, L.nx), dim3(64, 4), 0,
                           s, L, x, ld_x, H.z, H.p, qc, H.q, rc, H.r, sc);
    }
    FVX_LAUNCH_CHECK();
    return kOk;
  }
  for (int k = 0; k < n; ++k) {
    bool fused = false;
    int rc;
    if constexpr (std::is_same<T, float>::value) {
      if (pf) rc = pf(ctx, H.r, H.z, L.ny, L.nx, &sc->done, s);
      else rc = v_cycle_impl<T, KIND>(P, nullptr, H.r, H.z, H.az, H.q, sc, s, &fused);
    } else {
      rc = v_cycle_impl<T, KIND>(P, nullptr, H.r, H.z, H.az, H.q, sc, s, &fused);
    }
    if (rc) return rc;
    if (!fused) {
      const dim3 gs = grid_stencil(L.ny, L.nx);
      hipLaunchKernelGGL((pcg_az_dots_kernel<T, KIND>), gs, dim3(64, 4), 0, s, L, H.z, H.r, H.q,
                         H.az, P.partials, sc);
      hipLaunchKernelGGL((reduce_finish_kernel<kNumCgDots, kRedCG>), dim3(1),
                         dim3(kFinishThreads), 0, s, P.partials, (int)(gs.x * gs.y), sc, 0.0,
                         0.0);
    }
    bool recompute = false;
    if constexpr (std::is_same<T, float>::value)
      recompute = fused && H.lv.size() > 1 && wave_tiled<T, KIND>(L, &H.lv[1]);
    if (recompute) {
      if constexpr (std::is_same<T, float>::value)
        hipLaunchKernelGGL(pcg_cg_update_az_kernel<T>, grid_stencil(L.ny, L.nx), dim3(64, 4), 0, s, L,
                           x, ld_x, H.z, H.p, H.q, H.q, H.r, H.r, sc);
    } else {
      hipLaunchKernelGGL(pcg_cg_update_kernel<T>, grid_dir(L.ny, L.nx), dim3(64, 4), 0, s, L.ny,
                         L.nx, x, ld_x, H.z, H.az, H.p, H.q, H.r, sc);
    }
  }
  FVX_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
__global__ __launch_bounds__(256) void zero_pitched_kernel(int ny, int nx, T* __restrict__ x,
                                                           int64_t ld) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i < nx && j < ny) x[(int64_t)j * ld + i] = T(0);
}

// Enqueue PCG init + ``iters`` guarded iterations on slot ``slot``; no sync.
template <typename T>
int pcg_enqueue(Plan& P, const T* rhs, int64_t ld_rhs, T* x, int64_t ld_x, double rtol,
                double atol, bool use_x0, int iters, int slot, hipStream_t s) {
  Hier<T>& H = hier<T>(P);
  const auto& L = H.lv[0];
  PcgScalars* sc = &P.sc[2 * slot];
  if (!use_x0)
    hipLaunchKernelGGL(zero_pitched_kernel<T>, grid2(L.ny, L.nx), dim3(64, 4), 0, s, L.ny, L.nx,
                       x, ld_x);
  if (P.kind == kDirichlet)
    hipLaunchKernelGGL((pcg_init_kernel<T, kDirichlet>), grid2(L.ny, L.nx), dim3(64, 4), 0, s, L,
                       rhs, ld_rhs, use_x0 ? x : nullptr, ld_x, H.r, rtol * rtol, atol * atol,
                       P.partials, sc);
  else
    hipLaunchKernelGGL((pcg_init_kernel<T, kNoflux>), grid2(L.ny, L.nx), dim3(64, 4), 0, s, L,
                       rhs, ld_rhs, use_x0 ? x : nullptr, ld_x, H.r, rtol * rtol, atol * atol,
                       P.partials, sc);
  finish2<2, kRedInit>(P, L.ny, L.nx, sc, s, rtol * rtol, atol * atol);
  FVX_LAUNCH_CHECK();
  if (iters <= 0) return kOk;
  return (P.kind == kDirichlet) ? pcg_iters<T, kDirichlet>(P, H, x, ld_x, iters, sc, s)
                                : pcg_iters<T, kNoflux>(P, H, x, ld_x, iters, sc, s);
}

// FP64-accumulated true residual of x into the slot's outer block.
template <typename T>
int certify_enqueue(Plan& P, const T* rhs, int64_t ld_rhs, const T* x, int64_t ld_x, double rtol,
                    double atol, int slot, hipStream_t s) {
  Hier<T>& H = hier<T>(P);
  const auto& L = H.lv[0];
  PcgScalars* res = &P.sc[2 * slot + 1];
  if (P.kind == kDirichlet)
    hipLaunchKernelGGL((true_residual_kernel<T, kDirichlet>), grid2(L.ny, L.nx), dim3(64, 4), 0, s,
                       L, rhs, ld_rhs, x, ld_x, rtol * rtol, atol * atol, P.partials, res);
  else
    hipLaunchKernelGGL((true_residual_kernel<T, kNoflux>), grid2(L.ny, L.nx), dim3(64, 4), 0, s,
                       L, rhs, ld_rhs, x, ld_x, rtol * rtol, atol * atol, P.partials, res);
  finish2<2, kRedTrue>(P, L.ny, L.nx, res, s, rtol * rtol, atol * atol);
  FVX_LAUNCH_CHECK();
  return kOk;
}

// Synchronous PCG: predicted chunk, then single iterations until converged.
template <typename T>
int pcg_solve(Plan& P, const T* rhs, int64_t ld_rhs, T* x, int64_t ld_x, double rtol,
              double atol, int max_iters, bool use_x0, CgInfoC* info, int slot, hipStream_t s,
              double cert_rtol) {
  Hier<T>& H = hier<T>(P);
  PcgScalars* sc = &P.sc[2 * slot];
  int chunk = std::max(1, std::min(P.last_iters, max_iters));
  int rc = pcg_enqueue<T>(P, rhs, ld_rhs, x, ld_x, rtol, atol, use_x0, std::min(chunk, max_iters),
                          slot, s);
  if (rc) return rc;
  int done_iters = std::min(chunk, max_iters);
  while (true) {
    FVX_HIP(hipMemcpyAsync(&P.hsc[2 * slot], sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    FVX_HIP(hipStreamSynchronize(s));
    if (P.hsc[2 * slot].done || done_iters >= max_iters) break;
    rc = (P.kind == kDirichlet) ? pcg_iters<T, kDirichlet>(P, H, x, ld_x, 1, sc, s)
                                : pcg_iters<T, kNoflux>(P, H, x, ld_x, 1, sc, s);
    if (rc) return rc;
    done_iters += 1;
  }
  P.last_iters = std::max(1, P.hsc[2 * slot].iter);
  rc = certify_enqueue<T>(P, rhs, ld_rhs, x, ld_x, cert_rtol < 0 ? rtol : cert_rtol, atol, slot,
                          s);
  if (rc) return rc;
  FVX_HIP(hipMemcpyAsync(&P.hsc[2 * slot], &P.sc[2 * slot], 2 * sizeof(PcgScalars),
                         hipMemcpyDeviceToHost, s));
  FVX_HIP(hipStreamSynchronize(s));
  const PcgScalars& h = P.hsc[2 * slot];
  const PcgScalars& c = P.hsc[2 * slot + 1];
  if (info) {
    info->iterations = h.iter;
    info->outer = 0;
    info->converged = h.done && c.out_done;
    info->residual_norm = std::sqrt(c.out_rr);
    info->rhs_norm = std::sqrt(c.bb);
  }
  return kOk;
}

// FP64 refinement residual of x into the FP32 hierarchy's r (and e = 0), with
// its finishing reduction into the outer scalars ``out``.
template <typename TB>
// first: the residual starts a new solve and also clears its scalars.
static void launch_refine_residual(Plan& P, const TB* rhs, int64_t ld_rhs, const double* x,
                                   int64_t ld_x, double rtol, double atol, PcgScalars* out,
                                   hipStream_t s, bool reset_inner = false,
                                   double inner_rtol = 0.0, bool first = false) {
  Hier<float>& H = P.h32;
  const auto& L64 = P.h64.lv[0];
  const dim3 g = grid_res(L64.ny, L64.nx, P.kind);
  if (P.kind == kDirichlet)
    hipLaunchKernelGGL((refine_residual_kernel<kDirichlet, TB>), g, dim3(64, 4), 0, s, L64, rhs,
                       ld_rhs, x, ld_x, H.r, H.x, rtol * rtol, atol * atol, P.partials, out,
                       (int)first);
  else
    hipLaunchKernelGGL((refine_residual_kernel<kNoflux, TB>), g, dim3(64, 4), 0, s, L64, rhs,
                       ld_rhs, x, ld_x, H.r, H.x, rtol * rtol, atol * atol, P.partials, out,
                       (int)first);
  // with reset_inner, the same kernel resets the inner solve (out - 1) on
  // this residual
  if (reset_inner)
    hipLaunchKernelGGL((reduce_finish_kernel<2, kRedRefineReset>), dim3(1), dim3(kFinishThreads),
                       0, s, P.partials, (int)(g.x * g.y), out, rtol * rtol, atol * atol,
                       inner_rtol);
  else
    hipLaunchKernelGGL((reduce_finish_kernel<2, kRedRefine>), dim3(1), dim3(kFinishThreads), 0,
                       s, P.partials, (int)(g.x * g.y), out, rtol * rtol, atol * atol, 0.0);
}

// Enqueue the mixed-precision schedule: n_outer x [FP64 residual, FP32 PCG
// (n_inner[o] guarded iterations), FP64 update] and a closing FP64 residual.
// x += inner correction (with the inner solve's last PCG step when deferred).
static void launch_refine_update(Plan& P, double* x, int64_t ld_x, bool deferred,
                                 const PcgScalars* in, PcgScalars* out, hipStream_t s) {
  Hier<float>& H = P.h32;
  const auto& L32 = H.lv[0];
  if (deferred)
    hipLaunchKernelGGL(refine_update_step_kernel, grid_dir(L32.ny, L32.nx), dim3(64, 4), 0, s,
                       L32.ny, L32.nx, x, ld_x, H.x, H.z, H.p, in, out);
  else
    hipLaunchKernelGGL(refine_update_kernel, grid_dir(L32.ny, L32.nx), dim3(64, 4), 0, s, L32.ny,
                       L32.nx, x, ld_x, H.x, out);
}

double* plan_partials(Plan& P, size_t* capacity) {
  if (capacity) *capacity = P.npart;
  return P.partials;
}

void certify_finish(Plan& P, int nblk, double rtol, double atol, int slot, bool fp64_true,
                    hipStream_t s) {
  PcgScalars* out = &P.sc[2 * slot + 1];
  if (fp64_true)
    hipLaunchKernelGGL((reduce_finish_kernel<2, kRedTrue>), dim3(1), dim3(kFinishThreads), 0, s,
                       P.partials, nblk, out, rtol * rtol, atol * atol, 0.0);
  else
    hipLaunchKernelGGL((reduce_finish_kernel<2, kRedRefine>), dim3(1), dim3(kFinishThreads), 0,
                       s, P.partials, nblk, out, rtol * rtol, atol * atol, 0.0);
}

template <typename TB>
int mixed_enqueue(Plan& P, const TB* rhs, int64_t ld_rhs, double* x, int64_t ld_x, double rtol,
                  double atol, int n_outer, const int* n_inner, double inner_rtol, int slot,
                  hipStream_t s, PrecondFn pf, void* ctx, bool defer_final) {
  Hier<float>& H = P.h32;
  const auto& L64 = P.h64.lv[0];
  const auto& L32 = H.lv[0];
  PcgScalars* in = &P.sc[2 * slot];
  PcgScalars* out = &P.sc[2 * slot + 1];
  // out == in + 1: the residual's finishing kernel also resets the inner
  // solve; the first residual clears both scalar blocks of the slot
  auto residual = [&](bool reset_inner, bool first) {
    launch_refine_residual<TB>(P, rhs, ld_rhs, x, ld_x, rtol, atol, out, s, reset_inner,
                               inner_rtol, first);
  };
  for (int o = 0; o < n_outer; ++o) {
    residual(true, o == 0);
    bool deferred = false;
    int rc = (P.kind == kDirichlet)
                 ? pcg_iters<float, kDirichlet>(P, H, H.x, L32.nx, n_inner[o], in, s, pf, ctx,
                                                &deferred)
                 : pcg_iters<float, kNoflux>(P, H, H.x, L32.nx, n_inner[o], in, s, pf, ctx,
                                             &deferred);
    if (rc) return rc;
    launch_refine_update(P, x, ld_x, deferred, in, out, s);
  }
  if (n_outer == 0) hipLaunchKernelGGL(scalars_reset_kernel, dim3(1), dim3(64), 0, s, in, 2);
  if (!defer_final) residual(false, false);
  FVX_LAUNCH_CHECK();
  return kOk;
}

// Synchronous mixed solve: predicted schedule, then one more sweep at a time.
template <typename TB>
int mixed_solve(Plan& P, const TB* rhs, int64_t ld_rhs, double* x, int64_t ld_x, double rtol,
                double atol, int max_outer, int max_inner, double inner_rtol, CgInfoC* info,
                int slot, hipStream_t s, PrecondFn pf, void* ctx) {
  int fallbacks = 0;
  int inner = std::max(1, std::min(P.last_inner, max_inner));
  int outer_pred = std::max(1, std::min(P.last_outer, max_outer));
  int outer_total = 0;
  bool converged = false;
  int rc = 0;
  bool first = true;
  while (true) {
    const int no = first ? outer_pred : 1;
    if (first) {
      std::vector<int> sched(no, inner);
      rc = mixed_enqueue<TB>(P, rhs, ld_rhs, x, ld_x, rtol, atol, no, sched.data(), inner_rtol,
                             slot, s,
                             pf, ctx);
    } else {
      // Continue: do not reset the outer scalars (keeps the outer count).
      Hier<float>& H = P.h32;
      const auto& L32 = H.lv[0];
      PcgScalars* in = &P.sc[2 * slot];
      PcgScalars* out = &P.sc[2 * slot + 1];
      hipLaunchKernelGGL(inner_reset_kernel, dim3(1), dim3(1), 0, s, in, out, inner_rtol);
      bool deferred = false;
      rc = (P.kind == kDirichlet)
               ? pcg_iters<float, kDirichlet>(P, H, H.x, L32.nx, inner, in, s, pf, ctx, &deferred)
               : pcg_iters<float, kNoflux>(P, H, H.x, L32.nx, inner, in, s, pf, ctx, &deferred);
      if (rc) return rc;
      launch_refine_update(P, x, ld_x, deferred, in, out, s);
      launch_refine_residual<TB>(P, rhs, ld_rhs, x, ld_x, rtol, atol, out, s);
      FVX_LAUNCH_CHECK();
    }
    if (rc) return rc;
    first = false;
    FVX_HIP(hipMemcpyAsync(&P.hsc[2 * slot], &P.sc[2 * slot], 2 * sizeof(PcgScalars),
                           hipMemcpyDeviceToHost, s));
    FVX_HIP(hipStreamSynchronize(s));
    const PcgScalars& ho = P.hsc[2 * slot + 1];
    const PcgScalars& hi = P.hsc[2 * slot];
    outer_total = ho.outer;
    if (ho.out_done) {
      converged = true;
      break;
    }
    if (outer_total >= max_outer) break;
    if (!hi.done) {
      if (pf && inner >= max_inner) {
        // The learned inner solve hit its cap: fall back to the V-cycle.
        pf = nullptr;
        ctx = nullptr;
        ++fallbacks;
      }
      inner = std::min(max_inner, inner + 2);
    }
  }
  const PcgScalars& ho = P.hsc[2 * slot + 1];
  P.last_outer = std::max(1, ho.outer);
  P.last_inner = inner;
  if (info) {
    info->iterations = P.hsc[2 * slot].total;
    info->outer = ho.outer;
    info->converged = converged ? 1 : 0;
    info->residual_norm = std::sqrt(ho.out_rr);
    info->rhs_norm = std::sqrt(ho.bb);
    info->pad = fallbacks;
  }
  return kOk;
}

// Learned two-grid preconditioner on the FP32 hierarchy (see mg.hpp).
template <int KIND>
__global__ __launch_bounds__(256) void energy_dots_kernel(LevelDev<float> L,
                                                          const float* __restrict__ c,
                                                          const float* __restrict__ res,
                                                          double* partials,
                                                          const int* __restrict__ guard) {
  if (guard && *guard) return;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  double d0 = 0.0, d1 = 0.0;
  if (i < L.nx && j < L.ny) {
    const int64_t k = (int64_t)j * L.nx + i;
    const double cv = (double)c[k];
    d0 = (double)res[k] * cv;
    d1 = cv * (double)glob_Au<float, KIND>(L, c, j, i);
  }
  double v[2] = {d0, d1};
  block_partials<2, 256>(v, partials);
}

// alpha = <res, c> / <c, A c>: the multiple of c that minimises the energy
// norm of the error (A definite of either sign).
__global__ __launch_bounds__(kFinishThreads) void energy_finish_kernel(
    const double* __restrict__ partials, int nblk, double* alpha, const int* __restrict__ guard) {
  if (guard && *guard) return;
  __shared__ double red[kFinishThreads / 64][2];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double a0 = 0.0, a1 = 0.0;
#pragma unroll 8
  for (int b = tid; b < nblk; b += kFinishThreads) {
    a0 += partials[b];
    a1 += partials[(int64_t)nblk + b];
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  if (lane == 0) {
    red[wid][0] = a0;
    red[wid][1] = a1;
  }
  __syncthreads();
  if (tid == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int w = 0; w < kFinishThreads / 64; ++w) {
      s0 += red[w][0];
      s1 += red[w][1];
    }
    *alpha = s1 != 0.0 ? s0 / s1 : 0.0;
  }
}

__global__ __launch_bounds__(256) void add_scaled_kernel(int n, const double* __restrict__ alpha,
                                                         const float* __restrict__ c,
                                                         float* __restrict__ z,
                                                         const int* __restrict__ guard) {
  if (guard && *guard) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) z[k] += (float)(*alpha) * c[k];
}

float* fine_work(Plan& P, int k) {
  if (k < 0 || k >= 4) return nullptr;
  if (!P.work0[k]) {
    const auto& L = P.h32.lv[0];
    void* q = nullptr;
    if (hipMalloc(&q, (size_t)L.ny * L.nx * sizeof(float)) != hipSuccess) return nullptr;
    P.allocs.push_back(q);
    P.work0[k] = (float*)q;
  }
  if (!P.dscal) {
    void* q = nullptr;
    if (hipMalloc(&q, 8 * sizeof(double)) != hipSuccess) return nullptr;
    P.allocs.push_back(q);
    P.dscal = (double*)q;
  }
  return P.work0[k];
}

int learned_coarse_dims(const Plan& P, int* nyc, int* nxc) {
  if (!P.has32 || P.nl < 2) return kInvalidArgument;
  *nyc = P.h32.lv[1].ny;
  *nxc = P.h32.lv[1].nx;
  return kOk;
}

template <int KIND>
static int learned_two_grid_impl(Plan& P, CoarseNetFn fn, void* ctx, int sweeps, const float* r,
                                 float* z, const int* guard, hipStream_t s) {
  const auto& L = P.h32.lv[0];
  LevelDev<float> C = P.h32.lv[1];
  const int n = L.ny * L.nx;
  float* z0 = fine_work(P, 0);
  float* res = fine_work(P, 1);
  float* c = fine_work(P, 2);
  float* cw = fine_work(P, 3);
  if (!z0 || !res || !c || !cw) {
    set_error("allocation of the preconditioner work arrays failed");
    return kHipError;
  }
  float* rc = cw;
  float* zc = cw + (size_t)C.ny * C.nx;  // level 1 holds a quarter of level 0
  // z0 = S^sweeps(0; r), res = (r - A z0) m
  if (sweeps > 0) {
    smooth_chunks<float, KIND>(P, L, nullptr, true, r, sweeps, L.t, L.t2, z0, nullptr, s, true);
    hipLaunchKernelGGL((residual_kernel<float, KIND>), grid2(L.ny, L.nx), dim3(64, 4), 0, s, L,
                       z0, r, res);
  } else {
    hipLaunchKernelGGL(mask_copy_kernel<float>, grid2(L.ny, L.nx), dim3(64, 4), 0, s, L, nullptr,
                       z0, (int64_t)L.nx);
    hipLaunchKernelGGL(mask_copy_kernel<float>, grid2(L.ny, L.nx), dim3(64, 4), 0, s, L, r, res,
                       (int64_t)L.nx);
  }
  // coarse correction c = P net(R res), scaled by the energy-minimising alpha
  hipLaunchKernelGGL(restrict_kernel<float>, grid2(C.ny, C.nx), dim3(64, 4), 0, s, L, C, res, rc,
                     (int)(P.kind == kDirichlet));
  FVX_LAUNCH_CHECK();
  int rc_ = fn(ctx, rc, zc, C.ny, C.nx, C.mask, guard, s);
  if (rc_) return rc_;
  C.u = zc;
  hipLaunchKernelGGL((prolong_kernel<float, KIND>), grid2(L.ny, L.nx), dim3(64, 4), 0, s, L, C, c);
  hipLaunchKernelGGL(energy_dots_kernel<KIND>, grid2(L.ny, L.nx), dim3(64, 4), 0, s, L, c, res,
                     P.partials, guard);
  hipLaunchKernelGGL(energy_finish_kernel, dim3(1), dim3(kFinishThreads), 0, s, P.partials,
                     cdiv(L.nx, 64) * cdiv(L.ny, 4), P.dscal, guard);
  hipLaunchKernelGGL(add_scaled_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, P.dscal, c, z0,
                     guard);
  // post-smoothing from z0 into z
  if (sweeps > 0)
    smooth_chunks<float, KIND>(P, L, z0, false, r, sweeps, L.t, L.t2, z, nullptr, s, true);
  else
    hipLaunchKernelGGL(mask_copy_kernel<float>, grid2(L.ny, L.nx), dim3(64, 4), 0, s, L, z0, z,
                       (int64_t)L.nx);
  FVX_LAUNCH_CHECK();
  return kOk;
}

int learned_two_grid(Plan& P, CoarseNetFn fn, void* ctx, int sweeps, const float* r, float* z,
                     const int* guard, hipStream_t s) {
  if (!P.has32 || P.nl < 2) {
    set_error("the learned coarse correction needs an FP32 hierarchy with at least two levels");
    return kInvalidArgument;
  }
  return P.kind == kDirichlet
             ? learned_two_grid_impl<kDirichlet>(P, fn, ctx, sweeps, r, z, guard, s)
             : learned_two_grid_impl<kNoflux>(P, fn, ctx, sweeps, r, z, guard, s);
}

// Profiling hook support: a scalar block that never converges, so the
// guarded PCG kernels do their full work when launched back to back.
__global__ void bench_scalars_kernel(PcgScalars* sc) {
  sc->done = 0;
  sc->first = 0;
  sc->beta = 0.5;
  sc->alpha = 1e-6;
  sc->rz = 1.0;
  sc->tol2 = -1.0;
}

template <typename T, int KIND>
static int kernel_bench_impl(Plan& P, int which, int reps, hipStream_t s) {
  Hier<T>& H = hier<T>(P);
  const LevelDev<T>& L = H.lv[0];
  PcgScalars* sc = &P.sc[2 * (kSlots - 1)];
  if (reps == 0) {  // prepare: zero the vectors, arm the scalars
    for (T* v : {H.r, H.z, H.p, H.az, H.q, H.x, H.r2, H.q2})
      if (v)
        hipLaunchKernelGGL(zero_pitched_kernel<T>, grid2(L.ny, L.nx), dim3(64, 4), 0, s, L.ny,
                           L.nx, v, (int64_t)L.nx);
    if (P.nl >= 2) {
      const LevelDev<T>& C = H.lv[1];
      hipLaunchKernelGGL(zero_pitched_kernel<T>, grid2(C.ny, C.nx), dim3(64, 4), 0, s, C.ny, C.nx,
                         C.u, (int64_t)C.nx);
    }
    hipLaunchKernelGGL(bench_scalars_kernel, dim3(1), dim3(1), 0, s, sc);
    FVX_LAUNCH_CHECK();
    return kOk;
  }
  if (which >= 2 && P.nl < 2) {
    set_error("the level-0 V-cycle kernels need two levels");
    return kInvalidArgument;
  }
  bool step_fusable = false;
  if constexpr (std::is_same<T, float>::value && KIND == kDirichlet)
    step_fusable = wave_tiled<T, KIND>(L, &H.lv[1]) && P.n_pre <= kMaxFuse && H.r2 && H.q2;
  if (which == 3 && !step_fusable) {
    set_error("the down kernel with the PCG step needs an FP32 Dirichlet level 0");
    return kInvalidArgument;
  }
  for (int k = 0; k < reps; ++k) {
    if (which == 0) {
      hipLaunchKernelGGL((pcg_az_dots_kernel<T, KIND>), grid_stencil(L.ny, L.nx), dim3(64, 4), 0,
                         s, L, H.z, H.r, H.q, H.az, P.partials, sc);
    } else if (which == 1) {
      // the update kernel the PCG loop runs on this level
      bool recompute = false;
      if constexpr (std::is_same<T, float>::value)
        recompute = P.nl > 1 && wave_tiled<T, KIND>(L, &H.lv[1]);
      if (recompute) {
        if constexpr (std::is_same<T, float>::value)
          hipLaunchKernelGGL(pcg_cg_update_az_kernel<T>, grid_stencil(L.ny, L.nx), dim3(64, 4), 0, s,
                             L, H.x, (int64_t)L.nx, H.z, H.p, H.q, H.q, H.r, H.r, sc);
      } else {
        hipLaunchKernelGGL(pcg_cg_update_kernel<T>, grid_dir(L.ny, L.nx), dim3(64, 4), 0, s,
                           L.ny, L.nx, H.x, (int64_t)L.nx, H.z, H.az, H.p, H.q, H.r, sc);
      }
    } else if (which == 2) {
      // The last kernel of a preconditioning V-cycle on level 0: prolongation
      // of level 1's iterate, correction, two sweeps, a = A z, PCG dots.
      LevelDev<T> C = H.lv[1];
      launch_vcycle_up_cg<T, KIND>(P, L, C, H.p, H.r, H.z, H.q, H.az, sc, s);
    } else {
      // The first kernel of a preconditioning V-cycle on level 0 with the
      // previous PCG step applied first (r, q read; r2, q2 written).
      PcgUpdArgs<T> u{H.z, H.p, H.x, (int64_t)L.nx, H.q, H.q2, H.r, H.r2};
      const LevelDev<T>& C = H.lv[1];
      const int fz = P.n_pre < kMaxFuse ? P.n_pre : kMaxFuse;
      if (fz == 3) launch_tile<T, KIND, 3, kDownUpd, true>(P, L, &C, nullptr, H.r, L.u, nullptr,
                                                           sc, s, nullptr, true, &u);
      else if (fz == 2) launch_tile<T, KIND, 2, kDownUpd, true>(P, L, &C, nullptr, H.r, L.u,
                                                                nullptr, sc, s, nullptr, true, &u);
      else if (fz == 1) launch_tile<T, KIND, 1, kDownUpd, true>(P, L, &C, nullptr, H.r, L.u,
                                                                nullptr, sc, s, nullptr, true, &u);
      else launch_tile<T, KIND, 0, kDownUpd, true>(P, L, &C, nullptr, H.r, L.u, nullptr, sc, s,
                                                   nullptr, true, &u);
    }
  }
  FVX_LAUNCH_CHECK();
  return kOk;
}

const uint8_t* plan_mask0(const Plan& P) {
  return P.has32 ? P.h32.lv[0].mask : P.h64.lv[0].mask;
}
double plan_inv_diag0(const Plan& P) {
  return 1.0 / (2.0 / (P.dx * P.dx) + 2.0 / (P.dy * P.dy) + P.lam);
}

const PcgScalars* fetch_slots(Plan& P, int first_slot, int n, hipStream_t s) {
  if (hipMemcpyAsync(&P.hsc[2 * first_slot], &P.sc[2 * first_slot], 2 * n * sizeof(PcgScalars),
                     hipMemcpyDeviceToHost, s) != hipSuccess)
    return nullptr;
  if (hipStreamSynchronize(s) != hipSuccess) return nullptr;
  return &P.hsc[2 * first_slot];
}
int plan_set_wall(Plan& P, bool on) {
  if (!P.has32 || P.kind != kDirichlet) {
    set_error("the wall correction needs an FP32 Dirichlet hierarchy");
    return kInvalidArgument;
  }
  if (on) {
    // Every level that carries it must run on the wave-tile kernels: all
    // levels tiled down to a bottom level smoothed in tiles, sweeps fused
    // (at most 3 a side), and the bottom's later sweep chunks (non-zero
    // start) only on an unmasked bottom level.
    const Hier<float>& H = P.h32;
    const LevelDev<float>& B = H.lv[P.nl - 1];
    const bool tiled = P.nl == 1 || (P.lc == P.nl - 1 && (int64_t)B.ny * B.nx > kCoarseCells);
    const bool fused = P.n_pre <= kMaxFuse && P.n_post <= kMaxFuse;
    const bool bottom = P.n_coarse <= 3 || !B.mask;
    if (!(tiled && fused && bottom)) {
      set_error("the wall correction needs every level on the tiled kernels (bottom level over "
                "64x64 cells, at most 3 fused sweeps a side, 3 bottom sweeps on a masked bottom)");
      return kInvalidArgument;
    }
  }
  for (int l = 0; l < P.nl; ++l) {
    const double r = std::ldexp(1.0, l);
    P.h32.lv[l].wall_g = on ? (float)((1.0 - r) / (1.0 + r)) : 0.0f;
  }
  return kOk;
}
// Chebyshev weights of an m-sweep polynomial smoother on [lo, 2] (the
// spectrum of D^-1 A that the smoother has to damp): the reciprocals of the
// roots of the degree-m Chebyshev polynomial on that interval, ascending.
static void chebyshev_weights(double lo, int m, double* w) {
  for (int k = 0; k < m; ++k) {
    const double root =
        0.5 * (2.0 + lo) + 0.5 * (2.0 - lo) * std::cos(M_PI * (2 * k + 1) / (2.0 * m));
    w[k] = 1.0 / root;
  }
}

int plan_set_chebyshev(Plan& P, double lo) {
  if (lo <= 0.0) {
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < kMaxW; ++k) P.omw[r][k] = P.omega;
    P.cheb = false;
    return kOk;
  }
  if (lo >= 2.0 || P.nl < 2) {
    set_error("Chebyshev smoothing needs a hierarchy of 2 or more levels and 0 < lo < 2");
    return kInvalidArgument;
  }
  // The tiled kernels fuse up to 3 sweeps per launch (more restart the
  // weights); the bottom level runs on them (3 sweeps at most) or in the
  // coarse kernel with levels lc .. nl-1 (up to kMaxW weights).
  const int bny = P.has64 ? P.h64.lv[P.nl - 1].ny : P.h32.lv[P.nl - 1].ny;
  const int bnx = P.has64 ? P.h64.lv[P.nl - 1].nx : P.h32.lv[P.nl - 1].nx;
  const bool tiled_bottom = P.lc == P.nl - 1 && (int64_t)bny * bnx > kCoarseCells;
  const bool ok = P.n_pre >= 1 && P.n_pre <= kMaxFuse && P.n_post >= 1 && P.n_post <= kMaxFuse &&
                  P.n_coarse >= 1 && P.n_coarse <= (tiled_bottom ? kMaxFuse : kMaxW);
  if (!ok) {
    set_error("Chebyshev smoothing needs 1-3 sweeps before and after, and 1-3 bottom sweeps "
              "(1-8 in the coarse kernel)");
    return kInvalidArgument;
  }
  // pre-smoothing in ascending weight order, post-smoothing descending, the
  // bottom ascending (in exact arithmetic the order does not matter: the
  // sweeps' error factors commute)
  double w[kMaxW];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < kMaxW; ++k) P.omw[r][k] = P.omega;
  chebyshev_weights(lo, P.n_pre, w);
  for (int k = 0; k < P.n_pre; ++k) P.omw[0][k] = w[k];
  chebyshev_weights(lo, P.n_post, w);
  for (int k = 0; k < P.n_post; ++k) P.omw[1][k] = w[P.n_post - 1 - k];
  chebyshev_weights(lo, P.n_coarse, w);
  for (int k = 0; k < P.n_coarse; ++k) P.omw[2][k] = w[k];
  P.cheb = true;
  return kOk;
}
int plan_levels(const Plan& P) { return P.nl; }
bool plan_has_mixed(const Plan& P) { return P.has32 && P.has64; }
int plan_nx(const Plan& P) { return P.has64 ? P.h64.lv[0].nx : P.h32.lv[0].nx; }
int plan_ny(const Plan& P) { return P.has64 ? P.h64.lv[0].ny : P.h32.lv[0].ny; }
void plan_schedule(const Plan& P, int* outer, int* inner) {
  *outer = P.last_outer;
  *inner = P.last_inner;
}
void plan_set_schedule(Plan& P, int outer, int inner) {
  P.last_outer = std::max(1, outer);
  P.last_inner = std::max(1, inner);
}
int plan_max_inner(const Plan& P, int slot) {
  const PcgScalars& h = P.hsc[2 * slot];
  return std::max(h.max_inner, h.iter);
}
int plan_inner_iters(const Plan& P, int slot, int* per_outer, int n) {
  const PcgScalars& h = P.hsc[2 * slot];
  const PcgScalars& o = P.hsc[2 * slot + 1];
  const int no = std::min(o.outer, n);
  for (int k = 0; k < no; ++k) per_outer[k] = (k < 4) ? h.it_outer[k] : 0;
  // the last step is never followed by a reset: its count is still in iter
  if (no >= 1 && o.outer <= n) per_outer[no - 1] = h.iter;
  return no;
}
int plan_kind(const Plan& P) { return P.kind; }

template int pcg_enqueue<float>(Plan&, const float*, int64_t, float*, int64_t, double, double,
                                bool, int, int, hipStream_t);
template int pcg_enqueue<double>(Plan&, const double*, int64_t, double*, int64_t, double, double,
                                 bool, int, int, hipStream_t);
template int mixed_enqueue<float>(Plan&, const float*, int64_t, double*, int64_t, double, double,
                                  int, const int*, double, int, hipStream_t, PrecondFn, void*,
                                  bool);
template int mixed_enqueue<double>(Plan&, const double*, int64_t, double*, int64_t, double,
                                   double, int, const int*, double, int, hipStream_t, PrecondFn,
                                   void*, bool);
template int certify_enqueue<double>(Plan&, const double*, int64_t, const double*, int64_t,
                                     double, double, int, hipStream_t);
template int certify_enqueue<float>(Plan&, const float*, int64_t, const float*, int64_t, double,
                                    double, int, hipStream_t);
template int pcg_solve<double>(Plan&, const double*, int64_t, double*, int64_t, double, double,
                               int, bool, CgInfoC*, int, hipStream_t, double);
template int pcg_solve<float>(Plan&, const float*, int64_t, float*, int64_t, double, double, int,
                              bool, CgInfoC*, int, hipStream_t, double);
template int mixed_solve<float>(Plan&, const float*, int64_t, double*, int64_t, double, double,
                                int, int, double, CgInfoC*, int, hipStream_t, PrecondFn, void*);
template int mixed_solve<double>(Plan&, const double*, int64_t, double*, int64_t, double, double,
                                 int, int, double, CgInfoC*, int, hipStream_t, PrecondFn, void*);

}  // namespace fvx

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using fvx::Plan;

extern "C" {

typedef struct fvx_cg_info {
  int iterations;        // inner PCG iterations (summed over refinement sweeps)
  int outer;             // refinement sweeps (mixed precision), else 0
  int converged;         // 1 if the stopping rule was met
  int pad;
  double residual_norm;  // ||b - A x||_2 (FP64 true residual for mixed/FP64 solves)
  double rhs_norm;       // ||b||_2
} fvx_cg_info;

int fvx_mg_plan_create(const double* mask, const double* coeff, int ny, int nx, double dx,
                       double dy, double lambda_, int n_levels, int n_pre, int n_post,
                       int n_coarse, double omega, int kind, int precision, void** plan_out) {
  using namespace fvx;
  FVX_CHECK_ARG(mask && plan_out, "null argument");
  FVX_CHECK_ARG(ny >= 1 && nx >= 1, "empty grid");
  // The level kernels index cells (plus a halo row) with 32-bit offsets.
  FVX_CHECK_ARG((int64_t)(ny + 8) * (nx + 8) < ((int64_t)1 << 31),
                "grid too large: at most 2^31 cells including halo");
  FVX_CHECK_ARG(kind == kNoflux || kind == kDirichlet, "unknown operator kind %d", kind);
  FVX_CHECK_ARG(precision == 32 || precision == 64 || precision == 3264,
                "precision must be 32, 64 or 3264 (mixed)");
  FVX_CHECK_ARG(n_pre >= 0 && n_post >= 0 && n_coarse >= 0, "negative sweep count");
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  int nl = n_levels > 0 ? n_levels : compute_n_levels(ny, nx);
  FVX_CHECK_ARG(nl <= 16, "too many levels (%d)", nl);
  const int f = 1 << (nl - 1);
  if (ny % f != 0 || nx % f != 0) {
    set_error("Grid shape (%d, %d) is not divisible by 2^(n_levels-1) = %d.  Choose a different "
              "n_levels or pad the grid.", ny, nx, f);
    return kInvalidArgument;
  }
  Plan* P = new Plan();
  P->kind = kind;
  P->prec = precision;
  P->nl = nl;
  P->ny = ny;
  P->nx = nx;
  P->n_pre = n_pre;
  P->n_post = n_post;
  P->n_coarse = n_coarse;
  P->omega = omega;
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < kMaxW; ++k) P->omw[r][k] = omega;
  P->lam = lambda_;
  P->dx = dx;
  P->dy = dy;
  auto hl = build_host_levels(mask, coeff, ny, nx, dx, dy, lambda_, nl, kind);
  // First level small enough for the single-workgroup kernel (all below too).
  // (one-level hierarchies run entirely there; otherwise level 0 is tiled).
  coarse_lds_attributes();
  P->lc = nl;
  for (int l = 0; l < nl; ++l)
    if ((int64_t)hl[l].ny * hl[l].nx <= kCoarseCells) {
      P->lc = l;
      break;
    }
  // The bottom level is always the coarse kernel's (even when it is large).
  if (P->lc > nl - 1) P->lc = nl - 1;
  if (nl == 1) P->lc = 0;
  else if (P->lc == 0) P->lc = 1;
  int rc = 0;
  if (precision == 64 || precision == 3264) {
    rc = upload_hier<double>(*P, P->h64, hl);
    P->has64 = true;
  }
  if (!rc && (precision == 32 || precision == 3264)) {
    rc = upload_hier<float>(*P, P->h32, hl);
    P->has32 = true;
  }
  if (!rc) {
    // Partial sums: two per block of grid2 (residual, init, refinement and
    // energy dots), kNumCgDots per level-0 tile or stencil block (PCG dots).
    const size_t n2 = (size_t)2 * cdiv(nx, 64) * cdiv(ny, 4);
    const size_t ncg = (size_t)kNumCgDots * std::max<size_t>(
        (size_t)cdiv(nx, kTx) * cdiv(ny, kTy), (size_t)cdiv(nx, 256) * cdiv(ny, kStencilRows));
    const size_t npart = std::max(n2, ncg) + 64;
    P->npart = npart;
    hipError_t e = hipMalloc(&P->partials, npart * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&P->sc, 2 * kSlots * sizeof(PcgScalars));
    if (e == hipSuccess) e = hipMemset(P->sc, 0, 2 * kSlots * sizeof(PcgScalars));
    if (e == hipSuccess)
      e = hipHostMalloc(&P->hsc, 2 * kSlots * sizeof(PcgScalars), hipHostMallocDefault);
    if (e != hipSuccess) {
      set_error("HIP allocation failed: %s", hipGetErrorString(e));
      rc = kHipError;
    }
  }
  if (rc) {
    free_hier(P->h64);
    free_hier(P->h32);
    delete P;
    return rc;
  }
  *plan_out = P;
  return kOk;
}

int fvx_mg_plan_destroy(void* plan) {
  using namespace fvx;
  if (!plan) return kOk;
  Plan* P = (Plan*)plan;
  (void)hipDeviceSynchronize();
  free_hier(P->h64);
  free_hier(P->h32);
  (void)hipFree(P->partials);
  (void)hipFree(P->sc);
  (void)hipHostFree(P->hsc);
  delete P;
  return kOk;
}

// Chebyshev weights for the V-cycle's damped-Jacobi sweeps of a Dirichlet
// plan's FP32 hierarchy (smoothing interval [lo, 2] of D^-1 A); lo <= 0
// restores the uniform weight omega.
int fvx_mg_plan_set_chebyshev(void* plan, double lo) {
  using namespace fvx;
  FVX_CHECK_ARG(plan, "null plan");
  return plan_set_chebyshev(*(Plan*)plan, lo);
}

// Coarse-level wall correction of a mixed-precision Dirichlet plan's FP32
// hierarchy (on = 1) or off (0); see plan_set_wall.
int fvx_mg_plan_set_wall(void* plan, int on) {
  using namespace fvx;
  FVX_CHECK_ARG(plan, "null plan");
  return plan_set_wall(*(Plan*)plan, on != 0);
}

// Number of levels, level shapes (2 ints per level) and the coarse-kernel level.
int fvx_mg_plan_info(void* plan, int* n_levels, int* shapes, int* coarse_level) {
  using namespace fvx;
  FVX_CHECK_ARG(plan, "null plan");
  Plan* P = (Plan*)plan;
  if (n_levels) *n_levels = P->nl;
  if (coarse_level) *coarse_level = P->lc;
  if (shapes) {
    int ny = P->ny, nx = P->nx;
    for (int l = 0; l < P->nl; ++l) {
      shapes[2 * l] = ny;
      shapes[2 * l + 1] = nx;
      ny /= 2;
      nx /= 2;
    }
  }
  return kOk;
}

#define FVX_MG_T(SUF, T, HN)                                                                     \
  int fvx_mg_apply_operator_##SUF(void* plan, int level, const T* u, T* out, void* stream) {     \
    using namespace fvx;                                                                         \
    FVX_CHECK_ARG(plan && u && out, "null argument");                                            \
    Plan* P = (Plan*)plan;                                                                       \
    FVX_CHECK_ARG(P->has##HN, "plan was not built for this precision");                          \
    FVX_CHECK_ARG(level >= 0 && level < P->nl, "level %d out of range", level);                  \
    const auto& L = hier<T>(*P).lv[level];                                                       \
    FVX_CHECK_ARG(L.wall_g == T(0), "level %d carries the wall correction (solver-internal)",    \
                  level);                                                                        \
    if (P->kind == kDirichlet)                                                                   \
      hipLaunchKernelGGL((apply_op_kernel<T, kDirichlet>), grid2(L.ny, L.nx), dim3(64, 4), 0,    \
                         (hipStream_t)stream, L, u, out);                                        \
    else                                                                                         \
      hipLaunchKernelGGL((apply_op_kernel<T, kNoflux>), grid2(L.ny, L.nx), dim3(64, 4), 0,       \
                         (hipStream_t)stream, L, u, out);                                        \
    FVX_LAUNCH_CHECK();                                                                          \
    return kOk;                                                                                  \
  }                                                                                              \
  int fvx_mg_jacobi_##SUF(void* plan, int level, const T* u, const T* rhs, T* out, int n_iters,  \
                          void* stream) {                                                        \
    using namespace fvx;                                                                         \
    FVX_CHECK_ARG(plan && rhs && out, "null argument");                                          \
    Plan* P = (Plan*)plan;                                                                       \
    FVX_CHECK_ARG(P->has##HN, "plan was not built for this precision");                          \
    FVX_CHECK_ARG(level >= 0 && level < P->nl, "level %d out of range", level);                  \
    FVX_CHECK_ARG(n_iters >= 0, "negative n_iters");                                             \
    FVX_CHECK_ARG(out != u && out != rhs, "out must not alias inputs");                          \
    const auto& L = hier<T>(*P).lv[level];                                                       \
    FVX_CHECK_ARG(L.wall_g == T(0), "level %d carries the wall correction (solver-internal)",    \
                  level);                                                                        \
    FVX_CHECK_ARG(!P->cheb, "the plan smooths with Chebyshev weights (solver-internal)");        \
    hipStream_t s = (hipStream_t)stream;                                                         \
    if (n_iters == 0) {                                                                          \
      hipLaunchKernelGGL(mask_copy_kernel<T>, grid2(L.ny, L.nx), dim3(64, 4), 0, s, L, u, out,   \
                         (int64_t)L.nx);                                                         \
      hipLaunchKernelGGL(mask_copy_kernel<T>, grid2(L.ny, L.nx), dim3(64, 4), 0, s, L, out, out, \
                         (int64_t)L.nx);                                                         \
      FVX_LAUNCH_CHECK();                                                                        \
      return kOk;                                                                                \
    }                                                                                            \
    if (P->kind == kDirichlet)                                                                   \
      smooth_chunks<T, kDirichlet>(*P, L, u, u == nullptr, rhs, n_iters, L.t, L.t2, out,         \
                                   nullptr, s);                                                  \
    else                                                                                         \
      smooth_chunks<T, kNoflux>(*P, L, u, u == nullptr, rhs, n_iters, L.t, L.t2, out, nullptr,   \
                                s);                                                              \
    FVX_LAUNCH_CHECK();                                                                          \
    return kOk;                                                                                  \
  }                                                                                              \
  int fvx_mg_restrict_##SUF(void* plan, int level, const T* v, T* out, void* stream) {           \
    using namespace fvx;                                                                         \
    FVX_CHECK_ARG(plan && v && out, "null argument");                                            \
    Plan* P = (Plan*)plan;                                                                       \
    FVX_CHECK_ARG(P->has##HN, "plan was not built for this precision");                          \
    FVX_CHECK_ARG(level >= 0 && level + 1 < P->nl, "fine level %d has no coarse level", level);  \
    const auto& L = hier<T>(*P).lv[level];                                                       \
    const auto& C = hier<T>(*P).lv[level + 1];                                                   \
    hipLaunchKernelGGL(restrict_kernel<T>, grid2(C.ny, C.nx), dim3(64, 4), 0,                    \
                       (hipStream_t)stream, L, C, v, out, (int)(P->kind == kDirichlet));         \
    FVX_LAUNCH_CHECK();                                                                          \
    return kOk;                                                                                  \
  }                                                                                              \
  int fvx_mg_prolongate_##SUF(void* plan, int level, const T* vc, T* out, void* stream) {        \
    using namespace fvx;                                                                         \
    FVX_CHECK_ARG(plan && vc && out, "null argument");                                           \
    Plan* P = (Plan*)plan;                                                                       \
    FVX_CHECK_ARG(P->has##HN, "plan was not built for this precision");                          \
    FVX_CHECK_ARG(level >= 0 && level + 1 < P->nl, "fine level %d has no coarse level", level);  \
    const auto& L = hier<T>(*P).lv[level];                                                       \
    auto C = hier<T>(*P).lv[level + 1];                                                          \
    C.u = const_cast<T*>(vc);                                                                    \
    if (P->kind == kDirichlet)                                                                   \
      hipLaunchKernelGGL((prolong_kernel<T, kDirichlet>), grid2(L.ny, L.nx), dim3(64, 4), 0,     \
                         (hipStream_t)stream, L, C, out);                                        \
    else                                                                                         \
      hipLaunchKernelGGL((prolong_kernel<T, kNoflux>), grid2(L.ny, L.nx), dim3(64, 4), 0,        \
                         (hipStream_t)stream, L, C, out);                                        \
    FVX_LAUNCH_CHECK();                                                                          \
    return kOk;                                                                                  \
  }                                                                                              \
  int fvx_mg_v_cycle_##SUF(void* plan, const T* u, const T* rhs, T* out, void* stream) {         \
    using namespace fvx;                                                                         \
    FVX_CHECK_ARG(plan && rhs && out, "null argument");                                          \
    Plan* P = (Plan*)plan;                                                                       \
    FVX_CHECK_ARG(P->has##HN, "plan was not built for this precision");                          \
    FVX_CHECK_ARG(out != rhs && (u == nullptr || out != u), "out must not alias inputs");        \
    hipStream_t s = (hipStream_t)stream;                                                         \
    return v_cycle<T>(*P, u, rhs, out, nullptr, s);                                              \
  }                                                                                              \
  int fvx_mg_v_cycle_level_##SUF(void* plan, int level, const T* u, const T* rhs, T* out,       \
                                 void* stream) {                                                 \
    using namespace fvx;                                                                         \
    FVX_CHECK_ARG(plan && rhs && out, "null argument");                                          \
    Plan* P = (Plan*)plan;                                                                       \
    FVX_CHECK_ARG(P->has##HN, "plan was not built for this precision");                          \
    FVX_CHECK_ARG(level >= 0 && level < P->nl, "level %d outside 0..%d", level, P->nl - 1);      \
    FVX_CHECK_ARG(out != rhs && (u == nullptr || out != u), "out must not alias inputs");        \
    return v_cycle<T>(*P, u, rhs, out, nullptr, (hipStream_t)stream, level);                     \
  }                                                                                              \
  int fvx_mg_pcg_solve_##SUF(void* plan, const T* rhs, int64_t ld_rhs, T* x, int64_t ld_x,       \
                             double rtol, double atol, int max_iters, int use_x0,                \
                             fvx_cg_info* info, void* stream) {                                  \
    using namespace fvx;                                                                         \
    FVX_CHECK_ARG(plan && rhs && x, "null argument");                                            \
    Plan* P = (Plan*)plan;                                                                       \
    FVX_CHECK_ARG(P->has##HN, "plan was not built for this precision");                          \
    FVX_CHECK_ARG(max_iters >= 0, "negative max_iters");                                         \
    FVX_CHECK_ARG(ld_rhs >= P->nx && ld_x >= P->nx, "row pitch smaller than nx");               \
    return pcg_solve<T>(*P, rhs, ld_rhs, x, ld_x, rtol, atol, max_iters, use_x0 != 0,            \
                        (CgInfoC*)info, 0, (hipStream_t)stream);                                 \
  }

FVX_MG_T(f64, double, 64)
FVX_MG_T(f32, float, 32)
#undef FVX_MG_T

// FP32 multigrid-PCG inner solves with FP64 iterative refinement: the outer
// residual b - A x, the correction update and all inner products are FP64;
// the V-cycles, matvecs and vector updates of the inner solve are FP32.
// Converges on the FP64 true residual ||b - A x|| <= max(rtol ||b||, atol).
int fvx_mg_mixed_solve(void* plan, const double* rhs, int64_t ld_rhs, double* x, int64_t ld_x,
                       double rtol, double atol, int max_outer, int max_inner, double inner_rtol,
                       fvx_cg_info* info, void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(plan && rhs && x, "null argument");
  Plan* P = (Plan*)plan;
  FVX_CHECK_ARG(P->has64 && P->has32, "mixed solve needs a plan built with precision 3264");
  FVX_CHECK_ARG(max_outer >= 1 && max_inner >= 1, "iteration caps must be >= 1");
  FVX_CHECK_ARG(ld_rhs >= P->nx && ld_x >= P->nx, "row pitch smaller than nx");
  return mixed_solve<double>(*P, rhs, ld_rhs, x, ld_x, rtol, atol, max_outer, max_inner,
                             inner_rtol, (CgInfoC*)info, 0, (hipStream_t)stream);
}

// Profiling hook (no reference counterpart): launch ``reps`` back-to-back
// instances of one fine-level PCG kernel on the plan's own vectors with a
// scalar block that never converges.  which: 0 pcg_az_dots (a = A z and the
// PCG dots, the path of preconditioners other than the V-cycle),
// 1 pcg_cg_update (p = z + beta p, q = a + beta q, x += alpha p,
// r -= alpha q), 2 the level-0 up kernel of a preconditioning V-cycle
// (prolongation of level 1, correction, two sweeps, a = A z, PCG dots); bits
// 32 or 64 select the hierarchy.  reps = 0 prepares (zeroes the vectors, arms
// the scalars) and must precede timed calls.
int fvx_mg_kernel_bench(void* plan, int which, int bits, int reps, void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(plan, "null plan");
  Plan* P = (Plan*)plan;
  FVX_CHECK_ARG(which >= 0 && which <= 3,
                "which must be 0 (pcg_az_dots), 1 (pcg_cg_update), 2 (V-cycle up + dots) or "
                "3 (V-cycle down + PCG step)");
  FVX_CHECK_ARG((bits == 32 && P->has32) || (bits == 64 && P->has64),
                "plan has no hierarchy of that precision");
  FVX_CHECK_ARG(reps >= 0, "negative reps");
  hipStream_t s = (hipStream_t)stream;
  if (bits == 32)
    return P->kind == kDirichlet ? kernel_bench_impl<float, kDirichlet>(*P, which, reps, s)
                                 : kernel_bench_impl<float, kNoflux>(*P, which, reps, s);
  return P->kind == kDirichlet ? kernel_bench_impl<double, kDirichlet>(*P, which, reps, s)
                               : kernel_bench_impl<double, kNoflux>(*P, which, reps, s);
}

}  // extern "C"

// ===========================================================================
// Work-list batches of independent mixed-precision solves.
//
// An ensemble of QG members (uneven grid sizes, some with land masks, each
// its own Ld) or the vertical modes of a multi-layer model (one grid, one
// lambda per mode) hold one plan per problem.  Solving them one after the
// other would cost every member the full chain of dependent launches of its
// own solve, most of them latency-bound small levels.  A batch runs each
// operator of the mixed solve -- refinement residual, every V-cycle level's
// down and up kernel, the coarse solve, the PCG reductions and updates --
// as ONE launch over a work-list: entry e of a list is one member's argument
// block and the range of launch blocks that serve it; a block finds its entry
// by binary search over the ranges and runs the ordinary kernel body on the
// entry's tile.  Members whose hierarchy is shallower simply have no entry in
// the deeper positions.  Each member keeps its own device scalars, partials
// and convergence flags, so a converged member's blocks return at once while
// the others iterate.  Levels run on the LDS-tile kernels (masked and
// unmasked entries in separate lists, since the mask decides the kernel's
// LDS layout).
// ===========================================================================
namespace fvx {

// Entry owning launch block b: the last e with first[e] <= b.
__device__ __forceinline__ int wl_entry(const int* __restrict__ first, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

struct FinishArgs {
  const double* partials;
  int nblk;
  PcgScalars* sc;
};
struct UpdArgs {
  int ny, nx, gx;
  float *x, *z, *az, *p, *q, *r;
  int64_t ld_x;
  PcgScalars* sc;
};
struct ResArgs {
  LevelDev<double> L;
  const double* b;
  int64_t ld_b;
  const double* x;
  int64_t ld_x;
  float *r32, *e32;
  double* partials;
  PcgScalars* sc;
  int gx, gy;
};
struct RefUpdArgs {
  int ny, nx, gx;
  double* x;
  int64_t ld_x;
  const float* e;
  PcgScalars* sc;
  // the inner solve's last PCG step (fused batched solves): z, p and the
  // inner scalars (refine_update_step_body)
  const float* z;
  const float* p;
  const PcgScalars* in;
};

template <typename T, int KIND, int NS, int MODE, bool ZERO, bool MK>
__global__ __launch_bounds__(kTileThreads) void mg_tile_batch_kernel(
    const TileArgs<T>* __restrict__ args, const int* __restrict__ first, int n) {
  const int e = wl_entry(first, n, blockIdx.x);
  if (args[e].sc && args[e].sc->done) return;  // finished member: no loads
  mg_tile_body<T, KIND, NS, MODE, ZERO, MK>(args[e], blockIdx.x - first[e]);
}

// Work-list form of mg_wave_tile_kernel (FP32 Dirichlet levels of batched
// solves, with or without a land mask): the block's member by binary search
// of the block offsets, after the XCD-aware renumbering of the whole grid.
// (A persistent grid walking the list, stepping over finished members, ran
// at half the speed: each wave's tiles then run one after the other.)
template <int NS, int MODE, bool ZERO>
__global__ __launch_bounds__(256) void mg_wave_tile_batch_kernel(
    const TileArgs<float>* __restrict__ args, const int* __restrict__ first, int n) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int e = wl_entry(first, n, b);
  // A member whose solve has stopped loads nothing (the single-solve kernel
  // checks its flag after issuing the loads, to hide the flag's latency; in
  // a batch most blocks may belong to finished members).
  const TileArgs<float>& a = args[e];
  if (a.sc) {
    // (the down kernel with the PCG step still applies the step at done 2)
    const int f = a.sc->done;
    if (MODE == kDownUpd ? f == 1 : f != 0) return;
  }
  mg_wave_tile_body<NS, MODE, ZERO, kTileRows, true>(a, b - first[e], first[e + 1] - first[e]);
}

template <typename T, int KIND>
__global__ __launch_bounds__(kCoarseThreads) void mg_coarse_batch_kernel(
    const CoarseArgs<T>* __restrict__ args) {
  extern __shared__ __align__(16) unsigned char coarse_smem[];
  mg_coarse_lds_body<T, KIND>(args[blockIdx.x], coarse_smem);
}

template <int NV, int SITE>
__global__ __launch_bounds__(kFinishThreads) void finish_batch_kernel(
    const FinishArgs* __restrict__ args, double rtol2, double atol2, double inner_rtol) {
  const FinishArgs& a = args[blockIdx.x];
  reduce_finish_body<NV, SITE>(a.partials, a.nblk, a.sc, rtol2, atol2, inner_rtol);
}

__global__ __launch_bounds__(256) void pcg_update_batch_kernel(const UpdArgs* __restrict__ args,
                                                               const int* __restrict__ first,
                                                               int n) {
  const int e = wl_entry(first, n, blockIdx.x);
  const UpdArgs& a = args[e];
  const int lid = blockIdx.x - first[e];
  pcg_cg_update_body<float>(a.ny, a.nx, a.x, a.ld_x, a.z, a.az, a.p, a.q, a.r, a.sc,
                            lid % a.gx, lid / a.gx);
}

template <int KIND>
__global__ __launch_bounds__(256) void refine_residual_batch_kernel(
    const ResArgs* __restrict__ args, const int* __restrict__ first, int n, double rtol2,
    double atol2) {
  const int e = wl_entry(first, n, blockIdx.x);
  const ResArgs& a = args[e];
  const int lid = blockIdx.x - first[e];
  refine_residual_body<KIND, double>(a.L, a.b, a.ld_b, a.x, a.ld_x, a.r32, a.e32, rtol2, atol2,
                                     a.partials, a.sc, lid % a.gx, lid / a.gx, a.gx, a.gy);
}

__global__ __launch_bounds__(256) void refine_update_step_batch_kernel(
    const RefUpdArgs* __restrict__ args, const int* __restrict__ first, int n) {
  const int e = wl_entry(first, n, blockIdx.x);
  const RefUpdArgs& a = args[e];
  const int lid = blockIdx.x - first[e];
  refine_update_step_body(a.ny, a.nx, a.x, a.ld_x, a.e, a.z, a.p, a.in, a.sc, lid % a.gx,
                          lid / a.gx);
}

__global__ __launch_bounds__(256) void refine_update_batch_kernel(
    const RefUpdArgs* __restrict__ args, const int* __restrict__ first, int n) {
  const int e = wl_entry(first, n, blockIdx.x);
  const RefUpdArgs& a = args[e];
  const int lid = blockIdx.x - first[e];
  refine_update_body(a.ny, a.nx, a.x, a.ld_x, a.e, a.sc, lid % a.gx, lid / a.gx);
}

// Clears the scalar pairs of one slot of every member (block e: member e).
__global__ void scalars_reset_batch_kernel(PcgScalars* const* __restrict__ scs) {
  PcgScalars* sc = scs[blockIdx.x];
  const int k = threadIdx.x;
  if (k < 2) {
    PcgScalars& x = sc[k];
    x.rz = x.pq = x.rr = x.bb = x.tol2 = x.alpha = x.beta = x.zq = x.out_rr = x.out_tol2 = 0.0;
    x.iter = x.done = x.first = x.out_done = x.outer = x.total = x.max_inner = 0;
    x.it_outer[0] = x.it_outer[1] = x.it_outer[2] = x.it_outer[3] = 0;
  }
}

// Status gather: the (inner, outer) scalar pair of one slot of every member.
__global__ void scalars_gather_kernel(PcgScalars* const* __restrict__ scs, PcgScalars* out) {
  const int e = blockIdx.x, k = threadIdx.x;
  if (k < 2) out[2 * e + k] = scs[e][k];
}

// One work-list on the device: entry arguments and block ranges.
template <typename A>
struct DevList {
  A* args = nullptr;
  int* first = nullptr;
  int n = 0, blocks = 0;
};

template <typename A>
static int upload_list(const std::vector<A>& args, const std::vector<int>& nblk, DevList<A>& d,
                       std::vector<void*>& allocs) {
  d.n = (int)args.size();
  d.blocks = 0;
  if (d.n == 0) return kOk;
  std::vector<int> first(d.n + 1, 0);
  for (int e = 0; e < d.n; ++e) first[e + 1] = first[e] + nblk[e];
  d.blocks = first[d.n];
  FVX_HIP(hipMalloc(&d.args, sizeof(A) * d.n));
  allocs.push_back(d.args);
  FVX_HIP(hipMemcpy(d.args, args.data(), sizeof(A) * d.n, hipMemcpyHostToDevice));
  FVX_HIP(hipMalloc(&d.first, sizeof(int) * (d.n + 1)));
  allocs.push_back(d.first);
  FVX_HIP(hipMemcpy(d.first, first.data(), sizeof(int) * (d.n + 1), hipMemcpyHostToDevice));
  return kOk;
}

// The V-cycle and PCG work-lists of one scalar slot.
struct BatchSlot {
  bool built = false;
  // [MK][level] down and up lists of the LDS tile kernel, [level] lists of
  // the wave-tile kernel (FP32 Dirichlet levels); level 0's up lists are the
  // kUpCG ones
  std::vector<DevList<TileArgs<float>>> down[2], up[2], wdown, wup;
  // Fused PCG steps (every member's level 0 on wave tiles): level 0's down
  // kernel with the previous step, by parity of (r, q) -> (r2, q2), and the
  // level-0 up kernel with the dots reading (r, q) or (r2, q2)
  bool fused = false;
  DevList<TileArgs<float>> wdown_upd[2], wup_cg[2];
  DevList<CoarseArgs<float>> coarse;
  size_t coarse_lds = 0;
  DevList<FinishArgs> fin_cg, fin_res;
  DevList<UpdArgs> upd;
  DevList<RefUpdArgs> refupd;
  PcgScalars** scs_in = nullptr;  // device array of each member's inner scalars
  PcgScalars* status = nullptr;   // gathered pairs (device)
  PcgScalars* hstatus = nullptr;  // pinned host mirror
  // refinement residual and update lists by (rhs, x) array set (a list:
  // entries keep their addresses)
  struct ResEntry {
    std::vector<const void*> key;
    DevList<ResArgs> res;
    DevList<RefUpdArgs> ref;
  };
  std::list<ResEntry> res;
};

struct MgBatch {
  std::vector<Plan*> P;
  int kind = kDirichlet, n_pre = 0, n_post = 0, lc_max = 0;
  BatchSlot slot[kSlots];
  std::vector<void*> allocs;
};

static int batch_build_slot(MgBatch& B, int s) {
  BatchSlot& S = B.slot[s];
  if (S.built) return kOk;
  const int nm = (int)B.P.size();
  for (int mk = 0; mk < 2; ++mk) {
    S.down[mk].resize(B.lc_max);
    S.up[mk].resize(B.lc_max);
  }
  S.wdown.resize(B.lc_max);
  S.wup.resize(B.lc_max);
  // Wave tiles: 64 - 2 HL interior columns by kTileRows rows, four per block
  // (HL: the lane halo of the mode, wt_lane_halo).
  auto wave_tiles = [](TileArgs<float>& t, int hl) {
    t.tiles_x = cdiv(t.L.nx, 64 - 2 * hl);
    t.ntiles = t.tiles_x * cdiv(t.L.ny, kTileRows);
    return cdiv(t.ntiles, 4);
  };
  for (int l = 0; l < B.lc_max; ++l) {
    std::vector<TileArgs<float>> dn[2], upl[2], wdn, wupl;
    std::vector<int> dnb[2], upb[2], wdnb, wupb;
    for (int m = 0; m < nm; ++m) {
      Plan& Pm = *B.P[m];
      if (l >= Pm.lc) continue;
      Hier<float>& H = Pm.h32;
      PcgScalars* sc = &Pm.sc[2 * s];
      const LevelDev<float>& L = H.lv[l];
      const int mk = L.mask ? 1 : 0;
      TileArgs<float> a;
      std::memset(&a, 0, sizeof(a));
      a.L = L;
      a.C = H.lv[l + 1];
      a.omega = (float)Pm.omega;
      a.partials = Pm.partials;
      a.sc = sc;
      a.tiles_x = cdiv(L.nx, kTx);
      a.ntiles = a.tiles_x * cdiv(L.ny, kTy);
      // down: zero guess, b = (level 0: the inner residual r; else the
      // restricted residual), iterate into L.u, restriction into C.b
      const bool wave = B.kind == kDirichlet && wave_tiled<float, kDirichlet>(L, nullptr);
      TileArgs<float> d = a;
      for (int k = 0; k < 3; ++k) d.omw[k] = (float)Pm.omw[0][k];
      d.b = (l == 0) ? H.r : L.b;
      d.u_out = L.u;
      if (wave) {
        wdnb.push_back(wave_tiles(d, B.n_pre + 1));
        wdn.push_back(d);
      } else {
        dn[mk].push_back(d);
        dnb[mk].push_back(d.ntiles);
      }
      // up: coarse correction from level l+1's corrected iterate (the coarse
      // kernel's output at lc, else that level's up output L.t), post-sweeps
      TileArgs<float> u = a;
      for (int k = 0; k < 3; ++k) u.omw[k] = (float)Pm.omw[1][k];
      u.C.u = (l + 1 == Pm.lc) ? H.lv[l + 1].u : H.lv[l + 1].t;
      u.u_in = L.u;
      u.b = (l == 0) ? H.r : L.b;
      u.u_out = (l == 0) ? H.z : L.t;
      if (l == 0) {
        u.w = H.q;
        u.az_out = H.az;
      }
      if (wave) {
        wupb.push_back(wave_tiles(u, B.n_post + (l == 0 ? 1 : 0)));
        wupl.push_back(u);
      } else {
        upl[mk].push_back(u);
        upb[mk].push_back(u.ntiles);
      }
    }
    for (int mk = 0; mk < 2; ++mk) {
      if (int rc = upload_list(dn[mk], dnb[mk], S.down[mk][l], B.allocs)) return rc;
      if (int rc = upload_list(upl[mk], upb[mk], S.up[mk][l], B.allocs)) return rc;
    }
    if (int rc = upload_list(wdn, wdnb, S.wdown[l], B.allocs)) return rc;
    if (int rc = upload_list(wupl, wupb, S.wup[l], B.allocs)) return rc;
  }
  // Fused PCG steps when every member's level 0 runs on wave tiles with
  // fused sweeps (as the single solve's pcg_iters) and has at most 2048^2
  // cells.  Measured: the 96-member ensemble (members up to 512^2) 83.2 ->
  // 86.7 steps/s; the 3 x 4096^2 modal batch 30.6 -> 30.3, its fused level-0
  // down kernel taking 375 us against 259 us for the separate update and
  // down kernels once two of its three members have stopped.
  S.fused = B.kind == kDirichlet && B.n_pre <= kMaxFuse && B.n_post <= kMaxFuse && B.lc_max >= 1;
  for (int m = 0; m < nm && S.fused; ++m) {
    const Plan& Pm = *B.P[m];
    const LevelDev<float>& L0 = Pm.h32.lv[0];
    S.fused = Pm.lc >= 1 && Pm.nl > 1 && wave_tiled<float, kDirichlet>(L0, nullptr) &&
              (int64_t)L0.ny * L0.nx <= (int64_t)2048 * 2048 && Pm.h32.r2 && Pm.h32.q2;
  }
  if (S.fused) {
    for (int par = 0; par < 2; ++par) {
      std::vector<TileArgs<float>> dl, ul;
      std::vector<int> db, ubk;
      for (int m = 0; m < nm; ++m) {
        Plan& Pm = *B.P[m];
        Hier<float>& H = Pm.h32;
        const LevelDev<float>& L = H.lv[0];
        float* rc = par == 0 ? H.r : H.r2;
        float* qc = par == 0 ? H.q : H.q2;
        float* rn = par == 0 ? H.r2 : H.r;
        float* qn = par == 0 ? H.q2 : H.q;
        TileArgs<float> a;
        std::memset(&a, 0, sizeof(a));
        a.L = L;
        a.C = H.lv[1];
        a.omega = (float)Pm.omega;
        a.partials = Pm.partials;
        a.sc = &Pm.sc[2 * s];
        // down leg after the step that takes (rc, qc) to (rn, qn)
        TileArgs<float> d = a;
        for (int k = 0; k < 3; ++k) d.omw[k] = (float)Pm.omw[0][k];
        d.b = rc;
        d.u_out = L.u;
        d.upd = PcgUpdArgs<float>{H.z, H.p, H.x, (int64_t)L.nx, qc, qn, rc, rn};
        db.push_back(wave_tiles(d, B.n_pre + 2));
        dl.push_back(d);
        // up leg with the dots on (rc, qc): z only (the next down kernel
        // forms A z itself)
        TileArgs<float> u = a;
        for (int k = 0; k < 3; ++k) u.omw[k] = (float)Pm.omw[1][k];
        u.C.u = (1 == Pm.lc) ? H.lv[1].u : H.lv[1].t;
        u.u_in = L.u;
        u.b = rc;
        u.w = qc;
        u.u_out = H.z;
        ubk.push_back(wave_tiles(u, B.n_post + 1));
        ul.push_back(u);
      }
      if (int rc = upload_list(dl, db, S.wdown_upd[par], B.allocs)) return rc;
      if (int rc = upload_list(ul, ubk, S.wup_cg[par], B.allocs)) return rc;
    }
  }
  std::vector<CoarseArgs<float>> ca;
  std::vector<int> cb;
  std::vector<FinishArgs> fc, fr;
  std::vector<int> ones;
  std::vector<UpdArgs> ua;
  std::vector<int> ub;
  std::vector<RefUpdArgs> ra;
  std::vector<int> rb;
  std::vector<PcgScalars*> scs;
  for (int m = 0; m < nm; ++m) {
    Plan& Pm = *B.P[m];
    Hier<float>& H = Pm.h32;
    PcgScalars* sc = &Pm.sc[2 * s];
    scs.push_back(sc);
    if (Pm.lc < Pm.nl) {
      CoarseArgs<float> c;
      std::memset(&c, 0, sizeof(c));
      for (int l = 0; l < Pm.nl; ++l) c.lv[l] = H.lv[l];
      c.lc = Pm.lc;
      c.nl = Pm.nl;
      c.n_pre = Pm.n_pre;
      c.n_post = Pm.n_post;
      c.n_coarse = Pm.n_coarse;
      c.nonzero_init = 0;
      c.omega = (float)Pm.omega;
      coarse_weights(Pm, c);
      c.done = &sc->done;
      S.coarse_lds = std::max(S.coarse_lds, coarse_lds_bytes(c));
      ca.push_back(c);
      cb.push_back(1);
    }
    const LevelDev<float>& L0 = H.lv[0];
    // partials of the level-0 kUpCG kernel: one per LDS tile, or one per
    // block of four wave tiles
    int ncg = cdiv(L0.nx, kTx) * cdiv(L0.ny, kTy);
    if (B.kind == kDirichlet && wave_tiled<float, kDirichlet>(L0, nullptr)) {
      TileArgs<float> t;
      std::memset(&t, 0, sizeof(t));
      t.L = L0;
      ncg = wave_tiles(t, B.n_post + 1);
    }
    fc.push_back(FinishArgs{Pm.partials, ncg, sc});
    const dim3 gr = grid_res(Pm.h64.lv[0].ny, Pm.h64.lv[0].nx, Pm.kind);
    fr.push_back(FinishArgs{Pm.partials, (int)(gr.x * gr.y), sc + 1});
    ones.push_back(1);
    const dim3 gd = grid_dir(L0.ny, L0.nx);
    ua.push_back(UpdArgs{L0.ny, L0.nx, (int)gd.x, H.x, H.z, H.az, H.p, H.q, H.r, L0.nx, sc});
    ub.push_back((int)(gd.x * gd.y));
  }
  FVX_CHECK_ARG(S.coarse_lds <= kCoarseLdsMax, "a member's coarse levels exceed the LDS");
  if (int rc = upload_list(ca, cb, S.coarse, B.allocs)) return rc;
  if (int rc = upload_list(fc, ones, S.fin_cg, B.allocs)) return rc;
  if (int rc = upload_list(fr, ones, S.fin_res, B.allocs)) return rc;
  if (int rc = upload_list(ua, ub, S.upd, B.allocs)) return rc;
  FVX_HIP(hipMalloc(&S.scs_in, sizeof(PcgScalars*) * nm));
  B.allocs.push_back(S.scs_in);
  FVX_HIP(hipMemcpy(S.scs_in, scs.data(), sizeof(PcgScalars*) * nm, hipMemcpyHostToDevice));
  FVX_HIP(hipMalloc(&S.status, sizeof(PcgScalars) * 2 * nm));
  B.allocs.push_back(S.status);
  FVX_HIP(hipHostMalloc((void**)&S.hstatus, sizeof(PcgScalars) * 2 * nm, 0));
  S.built = true;
  return kOk;
}

// The refinement residual and update lists for one set of (rhs, x) arrays.
static int batch_res_list(MgBatch& B, int s, const double* const* rhs, const int64_t* ld_rhs,
                          double* const* x, const int64_t* ld_x, const DevList<ResArgs>** res,
                          const DevList<RefUpdArgs>** ref) {
  BatchSlot& S = B.slot[s];
  const int nm = (int)B.P.size();
  std::vector<const void*> key;
  for (int m = 0; m < nm; ++m) {
    key.push_back(rhs[m]);
    key.push_back(x[m]);
    key.push_back((const void*)(intptr_t)ld_rhs[m]);
    key.push_back((const void*)(intptr_t)ld_x[m]);
  }
  for (auto& e : S.res)
    if (e.key == key) {
      *res = &e.res;
      *ref = &e.ref;
      return kOk;
    }
  std::vector<ResArgs> ra;
  std::vector<RefUpdArgs> ua;
  std::vector<int> rb, ub;
  for (int m = 0; m < nm; ++m) {
    Plan& Pm = *B.P[m];
    const LevelDev<double>& L64 = Pm.h64.lv[0];
    const dim3 g = grid_res(L64.ny, L64.nx, Pm.kind);
    ResArgs a;
    std::memset(&a, 0, sizeof(a));
    a.L = L64;
    a.b = rhs[m];
    a.ld_b = ld_rhs[m];
    a.x = x[m];
    a.ld_x = ld_x[m];
    a.r32 = Pm.h32.r;
    a.e32 = Pm.h32.x;
    a.partials = Pm.partials;
    a.sc = &Pm.sc[2 * s + 1];
    a.gx = (int)g.x;
    a.gy = (int)g.y;
    ra.push_back(a);
    rb.push_back((int)(g.x * g.y));
    const LevelDev<float>& L0 = Pm.h32.lv[0];
    const dim3 gd = grid_dir(L0.ny, L0.nx);
    ua.push_back(RefUpdArgs{L0.ny, L0.nx, (int)gd.x, x[m], ld_x[m], Pm.h32.x, &Pm.sc[2 * s + 1],
                            Pm.h32.z, Pm.h32.p, &Pm.sc[2 * s]});
    ub.push_back((int)(gd.x * gd.y));
  }
  S.res.emplace_back();
  BatchSlot::ResEntry& e = S.res.back();
  e.key = key;
  if (int rc = upload_list(ra, rb, e.res, B.allocs)) return rc;
  if (int rc = upload_list(ua, ub, e.ref, B.allocs)) return rc;
  *res = &e.res;
  *ref = &e.ref;
  return kOk;
}

// step < 0: plain V-cycle (level 0 from zero on r, up leg with A z stored);
// step >= 0 (fused slots): level 0's down kernel first applies the previous
// PCG step of parity step - 1 (none for step 0), the up leg with the dots
// reads the (r, q) pair of parity step.
template <int KIND, int NPRE, int NPOST>
static void batch_vcycle(MgBatch& B, BatchSlot& S, hipStream_t s, int step = -1) {
  for (int l = 0; l < B.lc_max; ++l) {
    if constexpr (KIND == kDirichlet) {
      if (l == 0 && step > 0) {
        const DevList<TileArgs<float>>& w = S.wdown_upd[(step - 1) & 1];
        hipLaunchKernelGGL((mg_wave_tile_batch_kernel<NPRE, kDownUpd, true>), dim3(w.blocks),
                           dim3(256), 0, s, w.args, w.first, w.n);
        continue;
      }
    }
    const DevList<TileArgs<float>>& d0 = S.down[0][l];
    const DevList<TileArgs<float>>& d1 = S.down[1][l];
    if (d0.n)
      hipLaunchKernelGGL((mg_tile_batch_kernel<float, KIND, NPRE, kDown, true, false>),
                         dim3(d0.blocks), dim3(kTileThreads), 0, s, d0.args, d0.first, d0.n);
    if (d1.n)
      hipLaunchKernelGGL((mg_tile_batch_kernel<float, KIND, NPRE, kDown, true, true>),
                         dim3(d1.blocks), dim3(kTileThreads), 0, s, d1.args, d1.first, d1.n);
    if constexpr (KIND == kDirichlet) {
      const DevList<TileArgs<float>>& w = S.wdown[l];
      if (w.n)
        hipLaunchKernelGGL((mg_wave_tile_batch_kernel<NPRE, kDown, true>), dim3(w.blocks),
                           dim3(256), 0, s, w.args, w.first, w.n);
    }
  }
  if (S.coarse.n)
    hipLaunchKernelGGL((mg_coarse_batch_kernel<float, KIND>), dim3(S.coarse.n),
                       dim3(kCoarseThreads), S.coarse_lds, s, S.coarse.args);
  for (int l = B.lc_max - 1; l >= 0; --l) {
    const DevList<TileArgs<float>>& u0 = S.up[0][l];
    const DevList<TileArgs<float>>& u1 = S.up[1][l];
    if (l > 0) {
      if (u0.n)
        hipLaunchKernelGGL((mg_tile_batch_kernel<float, KIND, NPOST, kUp, false, false>),
                           dim3(u0.blocks), dim3(kTileThreads), 0, s, u0.args, u0.first, u0.n);
      if (u1.n)
        hipLaunchKernelGGL((mg_tile_batch_kernel<float, KIND, NPOST, kUp, false, true>),
                           dim3(u1.blocks), dim3(kTileThreads), 0, s, u1.args, u1.first, u1.n);
      if constexpr (KIND == kDirichlet) {
        const DevList<TileArgs<float>>& w = S.wup[l];
        if (w.n)
          hipLaunchKernelGGL((mg_wave_tile_batch_kernel<NPOST, kUp, false>), dim3(w.blocks),
                             dim3(256), 0, s, w.args, w.first, w.n);
      }
    } else {
      if (u0.n)
        hipLaunchKernelGGL((mg_tile_batch_kernel<float, KIND, NPOST, kUpCG, false, false>),
                           dim3(u0.blocks), dim3(kTileThreads), 0, s, u0.args, u0.first, u0.n);
      if (u1.n)
        hipLaunchKernelGGL((mg_tile_batch_kernel<float, KIND, NPOST, kUpCG, false, true>),
                           dim3(u1.blocks), dim3(kTileThreads), 0, s, u1.args, u1.first, u1.n);
      if constexpr (KIND == kDirichlet) {
        const DevList<TileArgs<float>>& w = step >= 0 ? S.wup_cg[step & 1] : S.wup[0];
        if (w.n)
          hipLaunchKernelGGL((mg_wave_tile_batch_kernel<NPOST, kUpCG, false>), dim3(w.blocks),
                             dim3(256), 0, s, w.args, w.first, w.n);
      }
    }
  }
}

template <int KIND, int NPRE, int NPOST>
static int batch_mixed_impl(MgBatch& B, int s, const DevList<ResArgs>& res,
                            const DevList<RefUpdArgs>& ref, double rtol, double atol,
                            int n_outer, const int* n_inner, double inner_rtol, hipStream_t st) {
  BatchSlot& S = B.slot[s];
  const int nm = (int)B.P.size();
  const double r2 = rtol * rtol, a2 = atol * atol;
  hipLaunchKernelGGL(scalars_reset_batch_kernel, dim3(nm), dim3(64), 0, st, S.scs_in);
  auto residual = [&](bool reset) {
    hipLaunchKernelGGL((refine_residual_batch_kernel<KIND>), dim3(res.blocks), dim3(64, 4), 0, st,
                       res.args, res.first, res.n, r2, a2);
    if (reset)
      hipLaunchKernelGGL((finish_batch_kernel<2, kRedRefineReset>), dim3(nm),
                         dim3(kFinishThreads), 0, st, S.fin_res.args, r2, a2, inner_rtol);
    else
      hipLaunchKernelGGL((finish_batch_kernel<2, kRedRefine>), dim3(nm), dim3(kFinishThreads), 0,
                         st, S.fin_res.args, r2, a2, 0.0);
  };
  for (int o = 0; o < n_outer; ++o) {
    residual(true);
    if (S.fused && n_inner[o] > 0) {
      // each PCG step is applied by the next V-cycle's level-0 down kernel;
      // the last one by the FP64 update
      for (int k = 0; k < n_inner[o]; ++k) {
        batch_vcycle<KIND, NPRE, NPOST>(B, S, st, k);
        hipLaunchKernelGGL((finish_batch_kernel<kNumCgDots, kRedCG>), dim3(nm),
                           dim3(kFinishThreads), 0, st, S.fin_cg.args, 0.0, 0.0, 0.0);
      }
      hipLaunchKernelGGL(refine_update_step_batch_kernel, dim3(ref.blocks), dim3(64, 4), 0, st,
                         ref.args, ref.first, ref.n);
      continue;
    }
    for (int k = 0; k < n_inner[o]; ++k) {
      batch_vcycle<KIND, NPRE, NPOST>(B, S, st);
      hipLaunchKernelGGL((finish_batch_kernel<kNumCgDots, kRedCG>), dim3(nm),
                         dim3(kFinishThreads), 0, st, S.fin_cg.args, 0.0, 0.0, 0.0);
      hipLaunchKernelGGL(pcg_update_batch_kernel, dim3(S.upd.blocks), dim3(64, 4), 0, st,
                         S.upd.args, S.upd.first, S.upd.n);
    }
    hipLaunchKernelGGL(refine_update_batch_kernel, dim3(ref.blocks), dim3(64, 4), 0, st, ref.args,
                       ref.first, ref.n);
  }
  residual(false);
  FVX_LAUNCH_CHECK();
  return kOk;
}

template <int KIND>
static int batch_mixed_dispatch(MgBatch& B, int s, const DevList<ResArgs>& res,
                                const DevList<RefUpdArgs>& ref, double rtol, double atol,
                                int n_outer, const int* n_inner, double inner_rtol,
                                hipStream_t st) {
#define FVX_BM(A, C)                                                                            \
  if (B.n_pre == A && B.n_post == C)                                                            \
    return batch_mixed_impl<KIND, A, C>(B, s, res, ref, rtol, atol, n_outer, n_inner, inner_rtol, \
                                        st);
  FVX_BM(1, 1) FVX_BM(2, 2) FVX_BM(3, 3) FVX_BM(1, 3) FVX_BM(2, 3) FVX_BM(3, 2)
#undef FVX_BM
  set_error("batched solves support V(1,1), V(2,2), V(3,3), V(1,3), V(2,3), V(3,2)");
  return kInvalidArgument;
}

}  // namespace fvx

extern "C" {

// A batch over existing mixed-precision plans (fvx_mg_plan_create with
// precision 3264): same operator kind and pre/post sweep counts, each plan's
// finest level tiled (more than 64 x 64 cells).  The plans must outlive the
// batch and must not run other solves on the batch's slots concurrently.
int fvx_mg_batch_create(void* const* plans, int n, void** batch_out) {
  using namespace fvx;
  FVX_CHECK_ARG(plans && n > 0 && batch_out, "null argument or empty batch");
  MgBatch* B = new MgBatch();
  for (int m = 0; m < n; ++m) {
    Plan* Pm = (Plan*)plans[m];
    if (!Pm || !Pm->has32 || !Pm->has64 || Pm->lc < 1) {
      delete B;
      set_error("member %d: needs a mixed-precision plan whose finest level is tiled", m);
      return kInvalidArgument;
    }
    if (m > 0 && (Pm->kind != B->kind || Pm->n_pre != B->n_pre || Pm->n_post != B->n_post)) {
      delete B;
      set_error("member %d: every member needs the same operator kind and sweep counts", m);
      return kInvalidArgument;
    }
    if (m == 0) {
      B->kind = Pm->kind;
      B->n_pre = Pm->n_pre;
      B->n_post = Pm->n_post;
    }
    B->lc_max = std::max(B->lc_max, Pm->lc);
    B->P.push_back(Pm);
  }
  coarse_lds_attributes();
  (void)hipFuncSetAttribute((const void*)mg_coarse_batch_kernel<float, kDirichlet>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCoarseLdsMax);
  (void)hipFuncSetAttribute((const void*)mg_coarse_batch_kernel<float, kNoflux>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCoarseLdsMax);
  *batch_out = B;
  return kOk;
}

int fvx_mg_batch_destroy(void* batch) {
  using namespace fvx;
  if (!batch) return kOk;
  MgBatch* B = (MgBatch*)batch;
  (void)hipDeviceSynchronize();
  for (void* p : B->allocs) (void)hipFree(p);
  for (int s = 0; s < kSlots; ++s)
    if (B->slot[s].hstatus) (void)hipHostFree(B->slot[s].hstatus);
  delete B;
  return kOk;
}

// Enqueue (no host synchronisation) the mixed-precision solve of every
// member on scalar slot ``slot``: rhs[m] / x[m] are member m's FP64 pitched
// right-hand side and warm start / solution; n_outer refinement steps with
// n_inner[o] inner PCG iterations each, device-guarded per member.  The first
// call with a given (slot, arrays) set builds its work-lists (synchronous
// uploads): make it before any graph capture that replays the call.
int fvx_mg_batch_mixed_enqueue(void* batch, const double* const* rhs, const int64_t* ld_rhs,
                               double* const* x, const int64_t* ld_x, double rtol, double atol,
                               int n_outer, const int* n_inner, double inner_rtol, int slot,
                               void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(batch && rhs && ld_rhs && x && ld_x && n_inner, "null argument");
  FVX_CHECK_ARG(slot >= 0 && slot < kSlots, "slot out of range");
  FVX_CHECK_ARG(n_outer >= 1 && n_outer <= 8, "n_outer must be in [1, 8]");
  MgBatch& B = *(MgBatch*)batch;
  if (int rc = batch_build_slot(B, slot)) return rc;
  const DevList<ResArgs>* res = nullptr;
  const DevList<RefUpdArgs>* ref = nullptr;
  if (int rc = batch_res_list(B, slot, rhs, ld_rhs, x, ld_x, &res, &ref)) return rc;
  hipStream_t s = (hipStream_t)stream;
  return B.kind == kDirichlet
             ? batch_mixed_dispatch<kDirichlet>(B, slot, *res, *ref, rtol, atol, n_outer,
                                                n_inner, inner_rtol, s)
             : batch_mixed_dispatch<kNoflux>(B, slot, *res, *ref, rtol, atol, n_outer, n_inner,
                                             inner_rtol, s);
}

// Per-member result of the last solve on ``slot`` (synchronises the stream):
// info[m] = {inner iterations, refinement steps, converged (FP64 residual),
// -, ||b - A x||, ||b||}; counts[m * 4 + o] = inner iterations of refinement
// step o (o < 4; the last step's when it was the last budgeted one).
int fvx_mg_batch_status(void* batch, int slot, int n_outer_budget, fvx_cg_info* info,
                        int* counts, void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(batch && info, "null argument");
  FVX_CHECK_ARG(slot >= 0 && slot < kSlots, "slot out of range");
  MgBatch& B = *(MgBatch*)batch;
  BatchSlot& S = B.slot[slot];
  FVX_CHECK_ARG(S.built, "no solve has been enqueued on this slot");
  const int nm = (int)B.P.size();
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(scalars_gather_kernel, dim3(nm), dim3(64), 0, s, S.scs_in, S.status);
  FVX_HIP(hipMemcpyAsync(S.hstatus, S.status, sizeof(PcgScalars) * 2 * nm,
                         hipMemcpyDeviceToHost, s));
  FVX_HIP(hipStreamSynchronize(s));
  for (int m = 0; m < nm; ++m) {
    const PcgScalars& in = S.hstatus[2 * m];
    const PcgScalars& out = S.hstatus[2 * m + 1];
    info[m].iterations = in.total;
    info[m].outer = out.outer;
    info[m].converged = out.out_done ? 1 : 0;
    info[m].pad = 0;
    info[m].residual_norm = std::sqrt(out.out_rr);
    info[m].rhs_norm = std::sqrt(out.bb);
    if (counts)
      for (int o = 0; o < 4; ++o)
        counts[m * 4 + o] = (o >= out.outer) ? 0
                            : (o == n_outer_budget - 1) ? in.iter
                                                        : in.it_outer[o];
  }
  return kOk;
}

}  // extern "C"
