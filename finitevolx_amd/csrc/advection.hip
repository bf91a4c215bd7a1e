// Fused flux-form advection tendency (Advection2D.__call__,
// advection/advection.py:329-420) for every reconstruction the drop-in
// exposes: naive, upwind1/2/3, weno3, wenoz3, weno5, wenoz5 and the four TVD
// limiters, unmasked or with per-face stencil tiers (advection/flux.py:100-253).
//
// One pass per tile: the tracer tile plus a 3-cell halo is staged in LDS from
// coalesced rows; every face flux of the tile is reconstructed exactly once
// (one upwind-selected evaluation per face, the reference evaluates both
// branches) and kept in LDS; the flux divergence, the mask and the zero
// 2-cell output ring are applied in the same kernel.  HBM traffic per cell:
// read h, u, v and write the tendency (+ 2 tier bytes + 1 mask byte masked).

#include <algorithm>
#include <cmath>
#include <mutex>

#include "fvx_common.hpp"

namespace fvx {

enum AdvMethod : int {
  kNaive = 0,
  kUpwind1 = 1,
  kUpwind2 = 2,
  kUpwind3 = 3,
  kWeno3 = 4,
  kWenoZ3 = 5,
  kWeno5 = 6,
  kWenoZ5 = 7,
  kTvd = 8,
};
enum Limiter : int { kMinmod = 0, kVanLeer = 1, kSuperbee = 2, kMC = 3 };

// --- point reconstructions (advection/weno.py) -----------------------------

template <typename T>
__device__ __forceinline__ T weno3_left(T qm, T q0, T qp) {  // weno.py:297-320
  const T eps = T(1e-8);
  const T qi1 = T(-0.5) * qm + T(1.5) * q0;
  const T qi2 = T(0.5) * (q0 + qp);
  const T b1 = (q0 - qm) * (q0 - qm), b2 = (qp - q0) * (qp - q0);
  const T t1 = b1 + eps, t2 = b2 + eps;
  const T w1 = T(1.0 / 3.0) / (t1 * t1), w2 = T(2.0 / 3.0) / (t2 * t2);
  return (w1 * qi1 + w2 * qi2) / (w1 + w2);
}

template <typename T>
__device__ __forceinline__ T weno3_right(T q0, T qp, T qpp) {  // weno.py:413-435
  const T eps = T(1e-8);
  const T qi1 = T(0.5) * (q0 + qp);
  const T qi2 = T(1.5) * qp - T(0.5) * qpp;
  const T b1 = (qp - q0) * (qp - q0), b2 = (qpp - qp) * (qpp - qp);
  const T t1 = b1 + eps, t2 = b2 + eps;
  const T w1 = T(2.0 / 3.0) / (t1 * t1), w2 = T(1.0 / 3.0) / (t2 * t2);
  return (w1 * qi1 + w2 * qi2) / (w1 + w2);
}

template <typename T>
__device__ __forceinline__ T wenoz3_left(T qm, T q0, T qp) {  // weno.py:323-347
  const T eps = T(1e-14);
  const T qi1 = T(-0.5) * qm + T(1.5) * q0;
  const T qi2 = T(0.5) * (q0 + qp);
  const T b1 = (q0 - qm) * (q0 - qm), b2 = (qp - q0) * (qp - q0);
  const T tau = fabs(b2 - b1);
  const T w1 = T(1.0 / 3.0) * (T(1) + tau / (b1 + eps));
  const T w2 = T(2.0 / 3.0) * (T(1) + tau / (b2 + eps));
  return (w1 * qi1 + w2 * qi2) / (w1 + w2);
}

template <typename T>
__device__ __forceinline__ T wenoz3_right(T q0, T qp, T qpp) {  // weno.py:438-459
  const T eps = T(1e-14);
  const T qi1 = T(0.5) * (q0 + qp);
  const T qi2 = T(1.5) * qp - T(0.5) * qpp;
  const T b1 = (qp - q0) * (qp - q0), b2 = (qpp - qp) * (qpp - qp);
  const T tau = fabs(b2 - b1);
  const T w1 = T(2.0 / 3.0) * (T(1) + tau / (b1 + eps));
  const T w2 = T(1.0 / 3.0) * (T(1) + tau / (b2 + eps));
  return (w1 * qi1 + w2 * qi2) / (w1 + w2);
}

// WENO5 on the upwind-ordered stencil (a, b, c, d, e) = (qmm, qm, q0, qp, qpp)
// for positive flow and (qppp, qpp, qp, q0, qm) for negative flow: the
// right-biased reference formula (weno.py:462-491) is the left-biased one
// (weno.py:350-376) on the mirrored stencil, so one evaluation serves both.
template <typename T, bool Z>
__device__ __forceinline__ T weno5_upwind(T a, T b, T c, T d, T e) {
  const T qi1 = T(1.0 / 3.0) * a - T(7.0 / 6.0) * b + T(11.0 / 6.0) * c;
  const T qi2 = T(-1.0 / 6.0) * b + T(5.0 / 6.0) * c + T(1.0 / 3.0) * d;
  const T qi3 = T(1.0 / 3.0) * c + T(5.0 / 6.0) * d - T(1.0 / 6.0) * e;
  const T k1 = T(13.0 / 12.0), k2 = T(0.25);
  const T s1 = a - T(2) * b + c, r1 = a - T(4) * b + T(3) * c;
  const T s2 = b - T(2) * c + d, r2 = b - d;
  const T s3 = c - T(2) * d + e, r3 = T(3) * c - T(4) * d + e;
  const T b1 = k1 * (s1 * s1) + k2 * (r1 * r1);
  const T b2 = k1 * (s2 * s2) + k2 * (r2 * r2);
  const T b3 = k1 * (s3 * s3) + k2 * (r3 * r3);
  T w1, w2, w3;
  if constexpr (!Z) {  // Jiang-Shu weights, eps 1e-8
    const T eps = T(1e-8);
    const T t1 = b1 + eps, t2 = b2 + eps, t3 = b3 + eps;
    w1 = T(0.1) / (t1 * t1);
    w2 = T(0.6) / (t2 * t2);
    w3 = T(0.3) / (t3 * t3);
  } else {  // WENO-Z (Borges et al. 2008), weno.py:379-410, eps 1e-16
    const T eps = T(1e-16);
    const T tau5 = fabs(b1 - b3);
    w1 = T(0.1) * (T(1) + tau5 / (b1 + eps));
    w2 = T(0.6) * (T(1) + tau5 / (b2 + eps));
    w3 = T(0.3) * (T(1) + tau5 / (b3 + eps));
  }
  return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3);
}

template <typename T, int L>
__device__ __forceinline__ T limiter(T r) {  // advection/limiters.py:34-123
  if constexpr (L == kMinmod) return fmax(T(0), fmin(T(1), r));
  if constexpr (L == kVanLeer) {
    const T ra = fabs(r);
    return (r + ra) / (T(1) + ra);
  }
  if constexpr (L == kSuperbee) return fmax(T(0), fmax(fmin(T(2) * r, T(1)), fmin(r, T(2))));
  if constexpr (L == kMC) return fmax(T(0), fmin(fmin((T(1) + r) / T(2), T(2) * r), T(2)));
  return T(0);
}

// Face value at the face between cells i and i+1 along one direction.
// s[k * st] is h[i + k] for k in [-2, 3]; ``fi`` is the face index and ``n``
// the array length along the direction; only faces 1..n-3 are evaluated
// (the ones the tendency reads, advection.py:410-417).
template <typename T, int M, int L>
__device__ __forceinline__ T face_value(const T* s, int st, int fi, int n, bool pos) {
  const T h0 = s[0], h1 = s[st];
  if constexpr (M == kNaive) return T(0.5) * (h0 + h1);
  if constexpr (M == kUpwind1) return pos ? h0 : h1;
  const T hm1 = s[-st], h2 = s[2 * st];
  if constexpr (M == kUpwind2)  // reconstruction.py:636-657
    return pos ? T(1.5) * h0 - T(0.5) * hm1 : T(1.5) * h1 - T(0.5) * h2;
  if constexpr (M == kUpwind3)  // reconstruction.py:682-711
    return pos ? T(-1.0 / 6.0) * hm1 + T(5.0 / 6.0) * h0 + T(1.0 / 3.0) * h1
               : T(1.0 / 3.0) * h0 + T(5.0 / 6.0) * h1 - T(1.0 / 6.0) * h2;
  if constexpr (M == kWeno3)  // reconstruction.py:744-764
    return pos ? weno3_left(hm1, h0, h1) : weno3_right(h0, h1, h2);
  if constexpr (M == kWenoZ3)  // reconstruction.py:788-808
    return pos ? wenoz3_left(hm1, h0, h1) : wenoz3_right(h0, h1, h2);
  if constexpr (M == kWeno5 || M == kWenoZ5) {  // reconstruction.py:832-864 / 900-932
    constexpr bool Z = (M == kWenoZ5);
    if (pos) {
      if (fi == 1) return Z ? wenoz3_left(hm1, h0, h1) : weno3_left(hm1, h0, h1);
      return weno5_upwind<T, Z>(s[-2 * st], hm1, h0, h1, h2);
    }
    if (fi == n - 3) return Z ? wenoz3_right(h0, h1, h2) : weno3_right(h0, h1, h2);
    return weno5_upwind<T, Z>(s[3 * st], h2, h1, h0, hm1);
  }
  if constexpr (M == kTvd) {  // reconstruction.py:1000-1050, _TVD_EPS = 1e-8
    const T eps = T(1e-8);
    if (pos) {
      const T dfwd = h1 - h0, dbwd = h0 - hm1;
      return h0 + T(0.5) * limiter<T, L>(dbwd / (dfwd + eps)) * dfwd;
    }
    const T dneg = h0 - h1, dneg2 = h1 - h2;
    return h1 + T(0.5) * limiter<T, L>(dneg2 / (dneg + eps)) * dneg;
  }
  return T(0);
}

// Masked dispatch (flux.py:231-250): the upwind cell's tier picks the
// reconstruction {2: upwind1, 4: the 3-point scheme, 6: the 5-point scheme};
// tier 0 (land / no stencil) gives a zero flux.
template <typename T, int M, int L, bool MASKED>
__device__ __forceinline__ T face_flux(const T* s, int st, int fi, int n, T vel, int tier) {
  const bool pos = vel >= T(0);
  if constexpr (!MASKED) {
    return face_value<T, M, L>(s, st, fi, n, pos) * vel;
  } else {
    constexpr int cap = (M == kWeno5 || M == kWenoZ5) ? 6 : 4;
    constexpr int M4 = (M == kWenoZ5) ? kWenoZ3 : (M == kTvd ? kTvd : kWeno3);
    const int t = tier < cap ? tier : cap;
    if (t == 2) return face_value<T, kUpwind1, L>(s, st, fi, n, pos) * vel;
    if (t == 4) return face_value<T, M4, L>(s, st, fi, n, pos) * vel;
    if (t == 6) return face_value<T, M, L>(s, st, fi, n, pos) * vel;
    return T(0);
  }
}

constexpr int kAdvTx = 64, kAdvTy = 8, kAdvHalo = 3, kAdvThreads = kAdvTx * kAdvTy;

// Optional fused Runge-Kutta stage: with hn set the kernel writes
// a hn + b h + cdt L(h) in place of L(h) (out may alias hn, not h).
template <typename T>
struct StageCoef {
  const T* hn = nullptr;
  double a = 0.0, b = 0.0, cdt = 0.0;
};

template <typename T, int M, int L, bool MASKED>
__global__ __launch_bounds__(kAdvThreads) void advection_kernel(
    const T* __restrict__ h, const T* __restrict__ u, const T* __restrict__ v,
    T* out, const uint8_t* __restrict__ tier_x, const uint8_t* __restrict__ tier_y,
    const uint8_t* __restrict__ mh, int ny, int nx, T idx, T idy, int tiles_x, int ntiles,
    const T* hn, T ca, T cb, T ccdt) {
  constexpr int SW = kAdvTx + 2 * kAdvHalo;
  __shared__ T sh[kAdvTy + 2 * kAdvHalo][SW];
  __shared__ T sfe[kAdvTy][kAdvTx + 1];
  __shared__ T sfn[kAdvTy + 1][kAdvTx];
  const int t = xcd_swizzle(blockIdx.x, ntiles);
  const int i0 = (t % tiles_x) * kAdvTx, j0 = (t / tiles_x) * kAdvTy;
  const int tid = threadIdx.x;

  for (int k = tid; k < (kAdvTy + 2 * kAdvHalo) * SW; k += kAdvThreads) {
    const int jj = k / SW, ii = k % SW;
    const int gj = j0 - kAdvHalo + jj, gi = i0 - kAdvHalo + ii;
    sh[jj][ii] = (gj >= 0 && gj < ny && gi >= 0 && gi < nx) ? h[(int64_t)gj * nx + gi] : T(0);
  }
  __syncthreads();

  // East faces fi = i0-1 .. i0+TX-1 on the tile's rows.
  for (int k = tid; k < kAdvTy * (kAdvTx + 1); k += kAdvThreads) {
    const int jj = k / (kAdvTx + 1), kk = k % (kAdvTx + 1);
    const int gj = j0 + jj, fi = i0 - 1 + kk;
    T f = T(0);
    if (gj >= 2 && gj <= ny - 3 && fi >= 1 && fi <= nx - 3) {
      const int64_t c = (int64_t)gj * nx + fi;
      const T vel = u[c];
      int tier = 0;
      if constexpr (MASKED) tier = (vel >= T(0)) ? tier_x[c] : tier_x[c + 1];
      f = face_flux<T, M, L, MASKED>(&sh[jj + kAdvHalo][kk + kAdvHalo - 1], 1, fi, nx, vel, tier);
    }
    sfe[jj][kk] = f;
  }
  // North faces fj = j0-1 .. j0+TY-1 on the tile's columns.
  for (int k = tid; k < (kAdvTy + 1) * kAdvTx; k += kAdvThreads) {
    const int kk = k / kAdvTx, ii = k % kAdvTx;
    const int gi = i0 + ii, fj = j0 - 1 + kk;
    T f = T(0);
    if (gi >= 2 && gi <= nx - 3 && fj >= 1 && fj <= ny - 3) {
      const int64_t c = (int64_t)fj * nx + gi;
      const T vel = v[c];
      int tier = 0;
      if constexpr (MASKED) tier = (vel >= T(0)) ? tier_y[c] : tier_y[c + nx];
      f = face_flux<T, M, L, MASKED>(&sh[kk + kAdvHalo - 1][ii + kAdvHalo], SW, fj, ny, vel, tier);
    }
    sfn[kk][ii] = f;
  }
  __syncthreads();

  const int tx = tid % kAdvTx, ty = tid / kAdvTx;
  const int gi = i0 + tx, gj = j0 + ty;
  if (gi < nx && gj < ny) {
    T r = T(0);
    if (gj >= 2 && gj <= ny - 3 && gi >= 2 && gi <= nx - 3)
      r = -((sfe[ty][tx + 1] - sfe[ty][tx]) * idx + (sfn[ty + 1][tx] - sfn[ty][tx]) * idy);
    const int64_t c = (int64_t)gj * nx + gi;
    if (mh) r = r * (T)(mh[c] ? 1 : 0);
    if (hn) r = ca * hn[c] + cb * sh[ty + kAdvHalo][tx + kAdvHalo] + ccdt * r;
    out[c] = r;
  }
}

template <typename T, int M, int L, bool MASKED>
static void launch_adv(const T* h, const T* u, const T* v, T* out, const uint8_t* tx_,
                       const uint8_t* ty_, const uint8_t* mh, int ny, int nx, double dx,
                       double dy, const StageCoef<T>& sc, hipStream_t s) {
  const int tiles_x = cdiv(nx, kAdvTx), tiles_y = cdiv(ny, kAdvTy);
  hipLaunchKernelGGL((advection_kernel<T, M, L, MASKED>), dim3(tiles_x * tiles_y),
                     dim3(kAdvThreads), 0, s, h, u, v, out, tx_, ty_, mh, ny, nx,
                     (T)(1.0 / dx), (T)(1.0 / dy), tiles_x, tiles_x * tiles_y, sc.hn, (T)sc.a,
                     (T)sc.b, (T)sc.cdt);
}

template <typename T, bool MASKED>
static int dispatch_adv(int method, int lim, const T* h, const T* u, const T* v, T* out,
                        const uint8_t* tx_, const uint8_t* ty_, const uint8_t* mh, int ny,
                        int nx, double dx, double dy, const StageCoef<T>& sc, hipStream_t s) {
#define FVX_ADV(M, L) \
  launch_adv<T, M, L, MASKED>(h, u, v, out, tx_, ty_, mh, ny, nx, dx, dy, sc, s)
  if constexpr (MASKED) {
    switch (method) {
      case kWeno3: FVX_ADV(kWeno3, 0); break;
      case kWeno5: FVX_ADV(kWeno5, 0); break;
      case kWenoZ5: FVX_ADV(kWenoZ5, 0); break;
      case kTvd:
        switch (lim) {
          case kMinmod: FVX_ADV(kTvd, kMinmod); break;
          case kVanLeer: FVX_ADV(kTvd, kVanLeer); break;
          case kSuperbee: FVX_ADV(kTvd, kSuperbee); break;
          case kMC: FVX_ADV(kTvd, kMC); break;
          default: set_error("unknown limiter %d", lim); return kInvalidArgument;
        }
        break;
      default:
        set_error("method %d does not support mask-based stencil dispatch", method);
        return kInvalidArgument;
    }
  } else {
    switch (method) {
      case kNaive: FVX_ADV(kNaive, 0); break;
      case kUpwind1: FVX_ADV(kUpwind1, 0); break;
      case kUpwind2: FVX_ADV(kUpwind2, 0); break;
      case kUpwind3: FVX_ADV(kUpwind3, 0); break;
      case kWeno3: FVX_ADV(kWeno3, 0); break;
      case kWenoZ3: FVX_ADV(kWenoZ3, 0); break;
      case kWeno5: FVX_ADV(kWeno5, 0); break;
      case kWenoZ5: FVX_ADV(kWenoZ5, 0); break;
      case kTvd:
        switch (lim) {
          case kMinmod: FVX_ADV(kTvd, kMinmod); break;
          case kVanLeer: FVX_ADV(kTvd, kVanLeer); break;
          case kSuperbee: FVX_ADV(kTvd, kSuperbee); break;
          case kMC: FVX_ADV(kTvd, kMC); break;
          default: set_error("unknown limiter %d", lim); return kInvalidArgument;
        }
        break;
      default: set_error("unknown advection method %d", method); return kInvalidArgument;
    }
  }
#undef FVX_ADV
  FVX_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int advection(int method, int lim, const T* h, const T* u, const T* v, T* out,
              const uint8_t* tier_x, const uint8_t* tier_y, const uint8_t* mh, int ny, int nx,
              double dx, double dy, hipStream_t s, const StageCoef<T>& sc = StageCoef<T>()) {
  FVX_CHECK_ARG(h && u && v && out, "null field pointer");
  FVX_CHECK_ARG(ny >= 5 && nx >= 5, "array too small (%d x %d); need >= 5 x 5", ny, nx);
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  FVX_CHECK_ARG((tier_x == nullptr) == (tier_y == nullptr), "tier_x and tier_y go together");
  FVX_CHECK_ARG(out != h, "out must not alias h (neighbouring tiles still read it)");
  if (tier_x)
    return dispatch_adv<T, true>(method, lim, h, u, v, out, tier_x, tier_y, mh, ny, nx, dx, dy,
                                 sc, s);
  return dispatch_adv<T, false>(method, lim, h, u, v, out, nullptr, nullptr, mh, ny, nx, dx, dy,
                                sc, s);
}

// ---------------------------------------------------------------------------
// Face fluxes of one direction (Reconstruction2D.<method>_x / _y and the
// *_masked variants, reconstruction.py:582-1307): every face 1..n-2 along the
// direction on rows (columns) 1..m-2, zero ring.  Faces the tendency never
// reads fall back as the reference does: at the last face n-2 positive flow
// takes the 3-point scheme in place of the 5-point one and negative flow
// takes upwind1 (the value beyond the face) for every scheme but naive.
// One thread per face; its six stencil values are loaded into registers
// (zero beyond the array).
// ---------------------------------------------------------------------------
template <typename T, int M, int L, bool MASKED>
__device__ __forceinline__ T face_flux_any(const T* s, int fi, int n, T vel, int tier) {
  if (fi == n - 2 && M != kNaive) {
    const bool pos = vel >= T(0);
    if (!pos) {
      if (MASKED && tier < 2) return T(0);
      return s[1] * vel;
    }
    if constexpr (M == kWeno5 || M == kWenoZ5) {
      constexpr int M3 = (M == kWeno5) ? kWeno3 : kWenoZ3;
      if (MASKED && tier < 2) return T(0);
      if (MASKED && tier == 2) return s[0] * vel;
      return face_value<T, M3, L>(s, 1, fi, n, true) * vel;
    }
  }
  return face_flux<T, M, L, MASKED>(s, 1, fi, n, vel, tier);
}

template <typename T, int M, int L, bool MASKED, bool YDIR>
__global__ __launch_bounds__(256) void face_flux_kernel(const T* __restrict__ h,
                                                        const T* __restrict__ vel,
                                                        T* __restrict__ out,
                                                        const uint8_t* __restrict__ tier, int ny,
                                                        int nx) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= nx || j >= ny) return;
  const int64_t c = (int64_t)j * nx + i;
  const int fi = YDIR ? j : i, n = YDIR ? ny : nx;
  const int oi = YDIR ? i : j, m = YDIR ? nx : ny;
  const int64_t st = YDIR ? nx : 1;
  T f = T(0);
  if (fi >= 1 && fi <= n - 2 && oi >= 1 && oi <= m - 2) {
    T w[6];
#pragma unroll
    for (int k = -2; k <= 3; ++k) {
      const int q = fi + k;
      w[k + 2] = (q >= 0 && q < n) ? h[c + k * st] : T(0);
    }
    const T v = vel[c];
    int t = 6;
    if constexpr (MASKED) t = (v >= T(0)) ? tier[c] : tier[c + st];
    f = face_flux_any<T, M, L, MASKED>(w + 2, fi, n, v, t);
  }
  out[c] = f;
}

template <typename T, int M, int L, bool MASKED>
static void launch_face_flux(int dir, const T* h, const T* vel, T* out, const uint8_t* tier,
                             int ny, int nx, hipStream_t s) {
  const dim3 g(cdiv(nx, 64), cdiv(ny, 4));
  if (dir == 0)
    hipLaunchKernelGGL((face_flux_kernel<T, M, L, MASKED, false>), g, dim3(256), 0, s, h, vel,
                       out, tier, ny, nx);
  else
    hipLaunchKernelGGL((face_flux_kernel<T, M, L, MASKED, true>), g, dim3(256), 0, s, h, vel, out,
                       tier, ny, nx);
}

template <typename T>
int face_flux_call(int method, int lim, int dir, const T* h, const T* vel, T* out,
                   const uint8_t* tier, int ny, int nx, hipStream_t s) {
  FVX_CHECK_ARG(h && vel && out, "null field pointer");
  FVX_CHECK_ARG(dir == 0 || dir == 1, "dir must be 0 (x, east faces) or 1 (y, north faces)");
  FVX_CHECK_ARG(ny >= 3 && nx >= 3, "array too small (%d x %d)", ny, nx);
  FVX_CHECK_ARG(lim >= 0 && lim <= 3, "unknown limiter %d", lim);
#define FVX_FF(M, L)                                                                 \
  (tier ? launch_face_flux<T, M, L, true>(dir, h, vel, out, tier, ny, nx, s)          \
        : launch_face_flux<T, M, L, false>(dir, h, vel, out, tier, ny, nx, s))
  if (tier) {
    FVX_CHECK_ARG(method == kWeno3 || method == kWeno5 || method == kWenoZ5 || method == kTvd,
                  "method %d does not support mask-based stencil dispatch", method);
  }
  switch (method) {
    case kNaive: launch_face_flux<T, kNaive, 0, false>(dir, h, vel, out, nullptr, ny, nx, s); break;
    case kUpwind1: launch_face_flux<T, kUpwind1, 0, false>(dir, h, vel, out, nullptr, ny, nx, s); break;
    case kUpwind2: launch_face_flux<T, kUpwind2, 0, false>(dir, h, vel, out, nullptr, ny, nx, s); break;
    case kUpwind3: launch_face_flux<T, kUpwind3, 0, false>(dir, h, vel, out, nullptr, ny, nx, s); break;
    case kWenoZ3: launch_face_flux<T, kWenoZ3, 0, false>(dir, h, vel, out, nullptr, ny, nx, s); break;
    case kWeno3: FVX_FF(kWeno3, 0); break;
    case kWeno5: FVX_FF(kWeno5, 0); break;
    case kWenoZ5: FVX_FF(kWenoZ5, 0); break;
    case kTvd:
      switch (lim) {
        case kMinmod: FVX_FF(kTvd, kMinmod); break;
        case kVanLeer: FVX_FF(kTvd, kVanLeer); break;
        case kSuperbee: FVX_FF(kTvd, kSuperbee); break;
        default: FVX_FF(kTvd, kMC); break;
      }
      break;
    default: set_error("unknown advection method %d", method); return kInvalidArgument;
  }
#undef FVX_FF
  FVX_LAUNCH_CHECK();
  return kOk;
}

// ---------------------------------------------------------------------------
// One rk3_ssp_step (timestepping/explicit_rk.py:74-121) of a doubly periodic
// tracer: per stage the periodic ghost fill of width ng (boundary.py:52-78,
// as the periodic advection tests do before each tendency), the Advection2D
// tendency and the Shu-Osher combination; h advances in place and leaves with
// its ghosts filled.  Scratch: the two stage states y1, y2.
// ---------------------------------------------------------------------------

// out = a x + b y + c z elementwise; out may alias x, y or z.
template <typename T>
__global__ __launch_bounds__(256) void rk_stage_kernel(int64_t n, T a, const T* x, T b,
                                                       const T* y, T c, const T* z, T* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += stride)
    out[k] = a * x[k] + b * y[k] + c * z[k];
}

}  // namespace fvx

extern "C" int fvx_enforce_periodic_f32(float*, int, int, int, void*);
extern "C" int fvx_enforce_periodic_f64(double*, int, int, int, void*);

namespace fvx {

static int periodic(float* f, int ny, int nx, int ng, hipStream_t s) {
  return fvx_enforce_periodic_f32(f, ny, nx, ng, s);
}
static int periodic(double* f, int ny, int nx, int ng, hipStream_t s) {
  return fvx_enforce_periodic_f64(f, ny, nx, ng, s);
}

template <typename T>
static int rk_stage(int64_t n, double a, const T* x, double b, const T* y, double c, const T* z,
                    T* out, hipStream_t s) {
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(rk_stage_kernel<T>, dim3(blocks), dim3(256), 0, s, n, (T)a, x, (T)b, y,
                     (T)c, z, out);
  FVX_LAUNCH_CHECK();
  return kOk;
}

// Argument checks shared by both entry points (no device work).
template <typename T>
static int adv_rk3_check(const T* h, const T* u, const T* v, int ny, int nx, int ng, double dx,
                         double dy, double dt) {
  FVX_CHECK_ARG(h && u && v, "null field pointer");
  FVX_CHECK_ARG(ng >= 1 && ny >= 3 * ng && nx >= 3 * ng && ny >= 5 && nx >= 5,
                "ghost width %d too large for %d x %d", ng, ny, nx);
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  FVX_CHECK_ARG(std::isfinite(dt), "dt must be finite");
  return kOk;
}

template <typename T>
static int adv_rk3_enqueue(int method, int lim, T* h, const T* u, const T* v,
                           const uint8_t* tier_x, const uint8_t* tier_y, const uint8_t* mh,
                           int ny, int nx, int ng, double dx, double dy, double dt, T* y1, T* y2,
                           hipStream_t s) {
  const int64_t n = (int64_t)ny * nx;
  auto L = [&](const T* q, T* k) {
    return advection<T>(method, lim, q, u, v, k, tier_x, tier_y, mh, ny, nx, dx, dy, s);
  };
  int rc;
  // k1 = L(h) into y1, y1 = h + dt k1
  if ((rc = periodic(h, ny, nx, ng, s)) || (rc = L(h, y1)) ||
      (rc = rk_stage<T>(n, 1.0, h, 0.0, h, dt, y1, y1, s)))
    return rc;
  // k2 = L(y1) into y2, y2 = 3/4 h + 1/4 y1 + 1/4 dt k2
  if ((rc = periodic(y1, ny, nx, ng, s)) || (rc = L(y1, y2)) ||
      (rc = rk_stage<T>(n, 0.75, h, 0.25, y1, 0.25 * dt, y2, y2, s)))
    return rc;
  // k3 = L(y2) into y1, h = 1/3 h + 2/3 y2 + 2/3 dt k3
  if ((rc = periodic(y2, ny, nx, ng, s)) || (rc = L(y2, y1)) ||
      (rc = rk_stage<T>(n, 1.0 / 3.0, h, 2.0 / 3.0, y2, 2.0 / 3.0 * dt, y1, h, s)))
    return rc;
  return periodic(h, ny, nx, ng, s);
}

// Stage arrays of the allocating entry point: one buffer per device, grown
// with hipMalloc when a call needs more and kept until the process ends.
static std::mutex g_stage_mu;
static void* g_stage_buf[64] = {};
static size_t g_stage_bytes[64] = {};

template <typename T>
static int adv_rk3_step(int method, int lim, T* h, const T* u, const T* v,
                        const uint8_t* tier_x, const uint8_t* tier_y, const uint8_t* mh, int ny,
                        int nx, int ng, double dx, double dy, double dt, hipStream_t s) {
  if (int rc = adv_rk3_check<T>(h, u, v, ny, nx, ng, dx, dy, dt)) return rc;
  const size_t bytes = 2 * (size_t)ny * nx * sizeof(T);
  int dev = 0;
  FVX_HIP(hipGetDevice(&dev));
  FVX_CHECK_ARG(dev >= 0 && dev < 64, "device index %d out of range", dev);
  T* buf = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_stage_mu);
    if (g_stage_bytes[dev] < bytes) {
      // earlier work on any stream may still use the old buffer
      FVX_HIP(hipDeviceSynchronize());
      if (g_stage_buf[dev]) FVX_HIP(hipFree(g_stage_buf[dev]));
      g_stage_buf[dev] = nullptr;
      g_stage_bytes[dev] = 0;
      FVX_HIP(hipMalloc(&g_stage_buf[dev], bytes));
      g_stage_bytes[dev] = bytes;
    }
    buf = (T*)g_stage_buf[dev];
  }
  return adv_rk3_enqueue<T>(method, lim, h, u, v, tier_x, tier_y, mh, ny, nx, ng, dx, dy, dt,
                            buf, buf + (size_t)ny * nx, s);
}

// The allocation-free variant keeps the two stage arrays in ``scratch``, a
// device array of kAdvScratchBytes that every GPU holds until the process
// ends; calls on one GPU share it (overlapping calls on one GPU race).  Host
// code never uses the array's name as an address: the first call on each
// device looks its address up with hipGetSymbolAddress (device current) and
// keeps it per device.
constexpr size_t kAdvScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kAdvScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64] = {};

static int scratch_ptr(void** out) {
  int dev = 0;
  FVX_HIP(hipGetDevice(&dev));
  FVX_CHECK_ARG(dev >= 0 && dev < 64, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return kHipError;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

template <typename T>
static int adv_rk3_step_noalloc(int method, int lim, T* h, const T* u, const T* v,
                                const uint8_t* tier_x, const uint8_t* tier_y,
                                const uint8_t* mh, int ny, int nx, int ng, double dx, double dy,
                                double dt, hipStream_t s) {
  if (int rc = adv_rk3_check<T>(h, u, v, ny, nx, ng, dx, dy, dt)) return rc;
  const size_t bytes = 2 * (size_t)ny * nx * sizeof(T);
  FVX_CHECK_ARG(bytes <= kAdvScratchBytes,
                "%d x %d needs %zu bytes of stage scratch; the allocation-free step holds "
                "%zu (two arrays of at most %zu elements)",
                ny, nx, bytes, kAdvScratchBytes, kAdvScratchBytes / (2 * sizeof(T)));
  void* p = nullptr;
  if (int rc = scratch_ptr(&p)) return rc;
  T* y1 = (T*)p;
  return adv_rk3_enqueue<T>(method, lim, h, u, v, tier_x, tier_y, mh, ny, nx, ng, dx, dy, dt, y1,
                            y1 + (size_t)ny * nx, s);
}

}  // namespace fvx

extern "C" {
#define FVX_ADV_RK3(SUF, T)                                                                      \
  int fvx_advection_rk3_step_##SUF(int method, int limiter, T* h, const T* u, const T* v,        \
                                   const uint8_t* tier_x, const uint8_t* tier_y,                \
                                   const uint8_t* mask_h, int ny, int nx, int ng, double dx,    \
                                   double dy, double dt, void* stream) {                        \
    return fvx::adv_rk3_step<T>(method, limiter, h, u, v, tier_x, tier_y, mask_h, ny, nx, ng,   \
                                dx, dy, dt, (hipStream_t)stream);                               \
  }                                                                                              \
  int fvx_advection_rk3_step_noalloc_##SUF(int method, int limiter, T* h, const T* u,           \
                                           const T* v, const uint8_t* tier_x,                   \
                                           const uint8_t* tier_y, const uint8_t* mask_h,        \
                                           int ny, int nx, int ng, double dx, double dy,        \
                                           double dt, void* stream) {                           \
    return fvx::adv_rk3_step_noalloc<T>(method, limiter, h, u, v, tier_x, tier_y, mask_h, ny,   \
                                        nx, ng, dx, dy, dt, (hipStream_t)stream);               \
  }
FVX_ADV_RK3(f32, float)
FVX_ADV_RK3(f64, double)
#undef FVX_ADV_RK3

int fvx_face_flux_f64(int method, int limiter, int dir, const double* h, const double* vel,
                      double* out, const uint8_t* tier, int ny, int nx, void* stream) {
  return fvx::face_flux_call<double>(method, limiter, dir, h, vel, out, tier, ny, nx,
                                     (hipStream_t)stream);
}
int fvx_face_flux_f32(int method, int limiter, int dir, const float* h, const float* vel,
                      float* out, const uint8_t* tier, int ny, int nx, void* stream) {
  return fvx::face_flux_call<float>(method, limiter, dir, h, vel, out, tier, ny, nx,
                                    (hipStream_t)stream);
}

#define FVX_ADV_STAGE(SUF, T)                                                                    \
  int fvx_advection_stage_##SUF(int method, int limiter, const T* hs, const T* u, const T* v,    \
                                const T* hn, T* out, const uint8_t* tier_x,                     \
                                const uint8_t* tier_y, const uint8_t* mask_h, int ny, int nx,   \
                                double dx, double dy, double a, double b, double cdt,           \
                                void* stream) {                                                 \
    FVX_CHECK_ARG(hn, "null hn");                                                               \
    fvx::StageCoef<T> sc;                                                                        \
    sc.hn = hn;                                                                                  \
    sc.a = a;                                                                                    \
    sc.b = b;                                                                                    \
    sc.cdt = cdt;                                                                                \
    return fvx::advection<T>(method, limiter, hs, u, v, out, tier_x, tier_y, mask_h, ny, nx, dx, \
                             dy, (hipStream_t)stream, sc);                                       \
  }
FVX_ADV_STAGE(f32, float)
FVX_ADV_STAGE(f64, double)
#undef FVX_ADV_STAGE

int fvx_advection2d_f64(int method, int limiter, const double* h, const double* u,
                        const double* v, double* out, const uint8_t* tier_x,
                        const uint8_t* tier_y, const uint8_t* mask_h, int ny, int nx,
                        double dx, double dy, void* stream) {
  return fvx::advection<double>(method, limiter, h, u, v, out, tier_x, tier_y, mask_h, ny, nx,
                                dx, dy, (hipStream_t)stream);
}
int fvx_advection2d_f32(int method, int limiter, const float* h, const float* u,
                        const float* v, float* out, const uint8_t* tier_x,
                        const uint8_t* tier_y, const uint8_t* mask_h, int ny, int nx,
                        double dx, double dy, void* stream) {
  return fvx::advection<float>(method, limiter, h, u, v, out, tier_x, tier_y, mask_h, ny, nx,
                               dx, dy, (hipStream_t)stream);
}
}
