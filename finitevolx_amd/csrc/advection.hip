// Fused flux-form advection tendency (Advection2D.__call__,
// advection/advection.py:329-420) for every reconstruction the drop-in
// exposes: naive, upwind1/2/3, weno3, wenoz3, weno5, wenoz5 and the four TVD
// limiters, unmasked or with per-face stencil tiers (advection/flux.py:100-253).
//
// One pass per wavefront tile (advection_kernel): the tracer column of each
// lane plus a 3-cell halo sits in registers, loaded from coalesced rows;
// every face flux of the tile is reconstructed exactly once (one
// upwind-selected, branch-free evaluation per face, the reference evaluates
// both branches); the flux divergence, the mask and the zero 2-cell output
// ring are applied in the same kernel.  HBM traffic per cell: read h, u, v
// and write the tendency (+ 2 tier bytes + 1 mask byte masked).

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
//
// The nonlinear weights are the reference's; only how they are divided out
// differs, because the divisions dominate the instruction count of these
// VALU-bound kernels.  FP32: each division is a product with the hardware
// reciprocal (v_rcp_f32, 1 ulp), inside the FP32 stencils' 1e-5 bound.  FP64:
// the normalised combination sum_k w_k q_k / sum_k w_k with w_k = d_k / t_k is
// rewritten over the common denominator prod_k t_k, so one IEEE division
// remains (the t_k are >= eps^2 and products of two or three of them stay in
// FP64's range); rounding differs from the reference's by a few ulps.

template <typename T>
__device__ __forceinline__ T rcp(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
  else return T(1) / x;
}

// (c1 q1 / t1 + c2 q2 / t2) / (c1 / t1 + c2 / t2)
template <typename T>
__device__ __forceinline__ T blend2(T c1, T t1, T q1, T c2, T t2, T q2) {
  if constexpr (sizeof(T) == 4) {
    const T w1 = c1 * rcp(t1), w2 = c2 * rcp(t2);
    return (w1 * q1 + w2 * q2) * rcp(w1 + w2);
  } else {
    const T a1 = c1 * t2, a2 = c2 * t1;
    return (a1 * q1 + a2 * q2) / (a1 + a2);
  }
}

// (c1 q1 / t1 + c2 q2 / t2 + c3 q3 / t3) / (c1 / t1 + c2 / t2 + c3 / t3)
template <typename T>
__device__ __forceinline__ T blend3(T c1, T t1, T q1, T c2, T t2, T q2, T c3, T t3, T q3) {
  if constexpr (sizeof(T) == 4) {
    const T w1 = c1 * rcp(t1), w2 = c2 * rcp(t2), w3 = c3 * rcp(t3);
    return (w1 * q1 + w2 * q2 + w3 * q3) * rcp(w1 + w2 + w3);
  } else {
    const T a1 = c1 * (t2 * t3), a2 = c2 * (t1 * t3), a3 = c3 * (t1 * t2);
    return (a1 * q1 + a2 * q2 + a3 * q3) / (a1 + a2 + a3);
  }
}

template <typename T>
__device__ __forceinline__ T weno3_left(T qm, T q0, T qp) {  // weno.py:297-320
  const T eps = T(1e-8);
  const T qi1 = T(-0.5) * qm + T(1.5) * q0;
  const T qi2 = T(0.5) * (q0 + qp);
  const T b1 = (q0 - qm) * (q0 - qm), b2 = (qp - q0) * (qp - q0);
  const T t1 = b1 + eps, t2 = b2 + eps;
  // w1 = (1/3) / t1^2, w2 = (2/3) / t2^2
  return blend2(T(1.0 / 3.0), t1 * t1, qi1, T(2.0 / 3.0), t2 * t2, qi2);
}

template <typename T>
__device__ __forceinline__ T weno3_right(T q0, T qp, T qpp) {  // weno.py:413-435
  const T eps = T(1e-8);
  const T qi1 = T(0.5) * (q0 + qp);
  const T qi2 = T(1.5) * qp - T(0.5) * qpp;
  const T b1 = (qp - q0) * (qp - q0), b2 = (qpp - qp) * (qpp - qp);
  const T t1 = b1 + eps, t2 = b2 + eps;
  return blend2(T(2.0 / 3.0), t1 * t1, qi1, T(1.0 / 3.0), t2 * t2, qi2);
}

// WENO-Z weights d_k (1 + tau / (b_k + eps)) = d_k (b_k + eps + tau) / (b_k + eps).
template <typename T>
__device__ __forceinline__ T wenoz3_left(T qm, T q0, T qp) {  // weno.py:323-347
  const T eps = T(1e-14);
  const T qi1 = T(-0.5) * qm + T(1.5) * q0;
  const T qi2 = T(0.5) * (q0 + qp);
  const T b1 = (q0 - qm) * (q0 - qm), b2 = (qp - q0) * (qp - q0);
  const T tau = fabs(b2 - b1);
  if constexpr (sizeof(T) == 4) {
    const T w1 = T(1.0 / 3.0) * (T(1) + tau * rcp(b1 + eps));
    const T w2 = T(2.0 / 3.0) * (T(1) + tau * rcp(b2 + eps));
    return (w1 * qi1 + w2 * qi2) * rcp(w1 + w2);
  } else {
    const T t1 = b1 + eps, t2 = b2 + eps;
    return blend2(T(1.0 / 3.0) * (t1 + tau), t1, qi1, T(2.0 / 3.0) * (t2 + tau), t2, qi2);
  }
}

template <typename T>
__device__ __forceinline__ T wenoz3_right(T q0, T qp, T qpp) {  // weno.py:438-459
  const T eps = T(1e-14);
  const T qi1 = T(0.5) * (q0 + qp);
  const T qi2 = T(1.5) * qp - T(0.5) * qpp;
  const T b1 = (qp - q0) * (qp - q0), b2 = (qpp - qp) * (qpp - qp);
  const T tau = fabs(b2 - b1);
  if constexpr (sizeof(T) == 4) {
    const T w1 = T(2.0 / 3.0) * (T(1) + tau * rcp(b1 + eps));
    const T w2 = T(1.0 / 3.0) * (T(1) + tau * rcp(b2 + eps));
    return (w1 * qi1 + w2 * qi2) * rcp(w1 + w2);
  } else {
    const T t1 = b1 + eps, t2 = b2 + eps;
    return blend2(T(2.0 / 3.0) * (t1 + tau), t1, qi1, T(1.0 / 3.0) * (t2 + tau), t2, qi2);
  }
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
  if constexpr (!Z) {  // Jiang-Shu weights d_k / (b_k + eps)^2, eps 1e-8
    const T eps = T(1e-8);
    const T t1 = b1 + eps, t2 = b2 + eps, t3 = b3 + eps;
    return blend3(T(0.1), t1 * t1, qi1, T(0.6), t2 * t2, qi2, T(0.3), t3 * t3, qi3);
  } else {  // WENO-Z (Borges et al. 2008), weno.py:379-410, eps 1e-16
    const T eps = T(1e-16);
    const T tau5 = fabs(b1 - b3);
    if constexpr (sizeof(T) == 4) {
      const T w1 = T(0.1) * (T(1) + tau5 * rcp(b1 + eps));
      const T w2 = T(0.6) * (T(1) + tau5 * rcp(b2 + eps));
      const T w3 = T(0.3) * (T(1) + tau5 * rcp(b3 + eps));
      return (w1 * qi1 + w2 * qi2 + w3 * qi3) * rcp(w1 + w2 + w3);
    } else {
      const T t1 = b1 + eps, t2 = b2 + eps, t3 = b3 + eps;
      return blend3(T(0.1) * (t1 + tau5), t1, qi1, T(0.6) * (t2 + tau5), t2, qi2,
                    T(0.3) * (t3 + tau5), t3, qi3);
    }
  }
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


// Optional fused Runge-Kutta stage: with hn set the kernel writes
// a hn + b h + cdt L(h) in place of L(h) (out may alias hn, not h).
template <typename T>
struct StageCoef {
  const T* hn = nullptr;
  double a = 0.0, b = 0.0, cdt = 0.0;
};

// Face value from the stencil mirrored to the upwind side of the face:
// m[k + 2] = h[i + k] for flow towards +x (k = -2 .. 2, the face between
// cells i and i+1) and h[i + 1 - k] for flow towards -x.  Every scheme's
// negative-flow formula is its positive-flow formula on the mirrored stencil
// (weno.py:413-491 against :297-410, reconstruction.py:636-1050), so one
// branch-free evaluation serves both signs.  ``near``: the face next to the
// array edge on the upwind side (face 1 for positive, n-3 for negative flow),
// where the 5-point schemes fall back to the 3-point ones.
template <typename T, int M, int L>
__device__ __forceinline__ T upwind_value(const T (&m)[5], bool near) {
  const T hm1 = m[1], h0 = m[2], h1 = m[3];
  if constexpr (M == kNaive) return T(0.5) * (h0 + h1);
  if constexpr (M == kUpwind1) return h0;
  if constexpr (M == kUpwind2) return T(1.5) * h0 - T(0.5) * hm1;
  if constexpr (M == kUpwind3)
    return T(-1.0 / 6.0) * hm1 + T(5.0 / 6.0) * h0 + T(1.0 / 3.0) * h1;
  if constexpr (M == kWeno3) return weno3_left(hm1, h0, h1);
  if constexpr (M == kWenoZ3) return wenoz3_left(hm1, h0, h1);
  if constexpr (M == kWeno5 || M == kWenoZ5) {
    constexpr bool Z = (M == kWenoZ5);
    if (near) return Z ? wenoz3_left(hm1, h0, h1) : weno3_left(hm1, h0, h1);
    return weno5_upwind<T, Z>(m[0], hm1, h0, h1, m[4]);
  }
  if constexpr (M == kTvd) {
    const T eps = T(1e-8);
    const T dfwd = h1 - h0, dbwd = h0 - hm1;
    return h0 + T(0.5) * limiter<T, L>(dbwd / (dfwd + eps)) * dfwd;
  }
  return T(0);
}

// The masked tier dispatch of face_flux on the mirrored stencil.
template <typename T, int M, int L, bool MASKED>
__device__ __forceinline__ T upwind_face(const T (&m)[5], bool near, int tier) {
  if constexpr (!MASKED) {
    return upwind_value<T, M, L>(m, near);
  } else {
    constexpr int cap = (M == kWeno5 || M == kWenoZ5) ? 6 : 4;
    constexpr int M4 = (M == kWenoZ5) ? kWenoZ3 : (M == kTvd ? kTvd : kWeno3);
    const int t = tier < cap ? tier : cap;
    if (t == 2) return upwind_value<T, kUpwind1, L>(m, near);
    if (t == 4) return upwind_value<T, M4, L>(m, near);
    if (t == 6) return upwind_value<T, M, L>(m, near);
    return T(0);
  }
}

// One wavefront per tile of 58 output columns (lanes 3..60; three halo lanes
// each side) x kAdvRows rows.  Each lane keeps its column of h on rows
// j0-3 .. j0+kAdvRows+2 in registers; the x stencils come from DPP lane
// shifts, the y stencils from neighbouring registers.  Each lane evaluates
// the east face of its cell (the west face is the east face of the lane
// beside it) and the north faces of its column, so every face flux is
// reconstructed once; no LDS, no barriers, no integer division per cell.
// Grids up to 512^2 take tiles of 2 rows: more wavefronts for the chip, a
// shorter chain of loads and arithmetic per wavefront.
constexpr int kAdvRows = 8, kAdvShortRows = 2, kAdvCols = 58;
constexpr int64_t kAdvShortMaxCells = 514 * 514;

template <typename T>
__device__ __forceinline__ T adv_e(T v) {
  return dpp_move<0x130>(v);  // wave_shl 1: value of lane + 1, 0 in lane 63
}
template <typename T>
__device__ __forceinline__ T adv_w(T v) {
  return dpp_move<0x138>(v);  // wave_shr 1: value of lane - 1, 0 in lane 0
}

template <typename T, int M, int L, bool MASKED, int R>
__global__ __launch_bounds__(256) void advection_kernel(
    const T* __restrict__ h, const T* __restrict__ u, const T* __restrict__ v,
    T* out, const uint8_t* __restrict__ tier_x, const uint8_t* __restrict__ tier_y,
    const uint8_t* __restrict__ mh, int ny, int nx, T idx, T idy, int tiles_x, int ntiles,
    const T* hn, T ca, T cb, T ccdt) {
  constexpr int HR = R + 6;
  const int lane = threadIdx.x & 63;
  const int tile =
      __builtin_amdgcn_readfirstlane(xcd_swizzle(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
  if (tile >= ntiles) return;  // whole wavefront (no block-level barrier below)
  const int tyi = tile / tiles_x;
  const int i0 = (tile - tyi * tiles_x) * kAdvCols, j0 = tyi * R;
  const int gi = i0 - 3 + lane;
  const bool colin = gi >= 0 && gi < nx;
  const int64_t gic = min(max(gi, 0), nx - 1);
  // h on rows j0-3+r (zero outside the array)
  T hr[HR];
#pragma unroll
  for (int r = 0; r < HR; ++r) {
    const int gj = j0 - 3 + r;
    const int64_t gjc = min(max(gj, 0), ny - 1);
    const T x = h[gjc * nx + gic];
    hr[r] = (colin && gj >= 0 && gj < ny) ? x : T(0);
  }
  // North faces fj = j0-1+t (between rows j0-1+t and j0+t), t = 0 .. R.
  T fn[R + 1];
#pragma unroll
  for (int t = 0; t <= R; ++t) {
    const int r = t + 2, fj = j0 - 1 + t;
    const int64_t fjc = min(max(fj, 0), ny - 1);
    const int64_t c = fjc * nx + gic;
    const T vel = v[c];
    const bool pos = vel >= T(0);
    const T m[5] = {pos ? hr[r - 2] : hr[r + 3], pos ? hr[r - 1] : hr[r + 2],
                    pos ? hr[r] : hr[r + 1], pos ? hr[r + 1] : hr[r],
                    pos ? hr[r + 2] : hr[r - 1]};
    int tier = 0;
    if constexpr (MASKED) tier = pos ? tier_y[c] : tier_y[min(fjc + 1, (int64_t)ny - 1) * nx + gic];
    const bool valid = colin && gi >= 2 && gi <= nx - 3 && fj >= 1 && fj <= ny - 3;
    const T f = upwind_face<T, M, L, MASKED>(m, pos ? fj == 1 : fj == ny - 3, tier) * vel;
    fn[t] = valid ? f : T(0);
  }
  const bool lout = colin && lane >= 3 && lane <= 60;
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int r = t + 3, gj = j0 + t;
    const int64_t gjc = min(gj, ny - 1);
    const int64_t c = gjc * nx + gic;
    // East face fi = gi of this row (between columns gi and gi+1).
    const T hc = hr[r];
    const T e1 = adv_e(hc), e2 = adv_e(e1), e3 = adv_e(e2);
    const T w1 = adv_w(hc), w2 = adv_w(w1);
    const T vel = u[c];
    const bool pos = vel >= T(0);
    const T m[5] = {pos ? w2 : e3, pos ? w1 : e2, pos ? hc : e1, pos ? e1 : hc, pos ? e2 : w1};
    int tier = 0;
    if constexpr (MASKED) {
      const int tx = tier_x[c];
      const int txe = __builtin_amdgcn_mov_dpp(tx, 0x130, 0xf, 0xf, true);  // lane + 1
      tier = pos ? tx : txe;
    }
    const bool valid = colin && gj >= 2 && gj <= ny - 3 && gi >= 1 && gi <= nx - 3;
    const T fe0 = upwind_face<T, M, L, MASKED>(m, pos ? gi == 1 : gi == nx - 3, tier) * vel;
    const T fe = valid ? fe0 : T(0);
    const T fw = adv_w(fe);
    if (lout && gj < ny) {
      T res = T(0);
      if (gj >= 2 && gj <= ny - 3 && gi >= 2 && gi <= nx - 3)
        res = -((fe - fw) * idx + (fn[t + 1] - fn[t]) * idy);
      if (mh) res = res * (T)(mh[c] ? 1 : 0);
      if (hn) res = ca * hn[c] + cb * hc + ccdt * res;
      out[c] = res;
    }
  }
}

template <typename T, int M, int L, bool MASKED>
static void launch_adv(const T* h, const T* u, const T* v, T* out, const uint8_t* tx_,
                       const uint8_t* ty_, const uint8_t* mh, int ny, int nx, double dx,
                       double dy, const StageCoef<T>& sc, hipStream_t s) {
  const bool short_tiles = (int64_t)ny * nx <= kAdvShortMaxCells;
  const int rows = short_tiles ? kAdvShortRows : kAdvRows;
  const int tiles_x = cdiv(nx, kAdvCols), tiles_y = cdiv(ny, rows);
  const int ntiles = tiles_x * tiles_y;
  if (short_tiles)
    hipLaunchKernelGGL((advection_kernel<T, M, L, MASKED, kAdvShortRows>), dim3(cdiv(ntiles, 4)),
                       dim3(256), 0, s, h, u, v, out, tx_, ty_, mh, ny, nx, (T)(1.0 / dx),
                       (T)(1.0 / dy), tiles_x, ntiles, sc.hn, (T)sc.a, (T)sc.b, (T)sc.cdt);
  else
    hipLaunchKernelGGL((advection_kernel<T, M, L, MASKED, kAdvRows>), dim3(cdiv(ntiles, 4)),
                       dim3(256), 0, s, h, u, v, out, tx_, ty_, mh, ny, nx, (T)(1.0 / dx),
                       (T)(1.0 / dy), tiles_x, ntiles, sc.hn, (T)sc.a, (T)sc.b, (T)sc.cdt);
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
// as the periodic advection tests do before each tendency) and the
// Advection2D tendency with the Shu-Osher combination fused into its launch;
// h advances in place and leaves with its ghosts filled.  Scratch: the two
// stage states y1, y2.
// ---------------------------------------------------------------------------

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
  // Each stage is one advection launch with the Shu-Osher combination fused
  // (out = a h + b q + cdt L(q) on every cell, q the stage's input).
  auto stage = [&](const T* q, T* out, double a, double b, double cdt) {
    StageCoef<T> sc;
    sc.hn = h;
    sc.a = a;
    sc.b = b;
    sc.cdt = cdt;
    return advection<T>(method, lim, q, u, v, out, tier_x, tier_y, mh, ny, nx, dx, dy, s, sc);
  };
  int rc;
  // y1 = h + dt L(h)
  if ((rc = periodic(h, ny, nx, ng, s)) || (rc = stage(h, y1, 0.0, 1.0, dt))) return rc;
  // y2 = 3/4 h + 1/4 y1 + 1/4 dt L(y1)
  if ((rc = periodic(y1, ny, nx, ng, s)) || (rc = stage(y1, y2, 0.75, 0.25, 0.25 * dt))) return rc;
  // h = 1/3 h + 2/3 y2 + 2/3 dt L(y2) (the output may alias h: read per cell)
  if ((rc = periodic(y2, ny, nx, ng, s)) ||
      (rc = stage(y2, h, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt)))
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
