// 1.5-layer quasi-geostrophic step (build-defined composition of reference
// operators; nearest reference: scripts/qg_1p5_layer.py:420-497):
//
//   dq/dt = -J(psi, q + beta (y - y0)) + F(y) - nu4 D(D(q))
//   (lap - 1/Ld^2) psi = q,  psi = 0 on land and beyond the basin walls
//
// with J = arakawa_jacobian (operators/jacobian.py:22-131), D the no-flux
// Diffusion2D Laplacian (BiharmonicDiffusion2D, diffusion/diffusion.py:543-574),
// F = -F0 sin(2 pi y / Ly) (qg_1p5_layer.py:363) and rk3_ssp_step stepping
// (timestepping/explicit_rk.py:104-121).
//
// Per RK stage: one multigrid-PCG Helmholtz solve (FP64, or FP32 with FP64
// refinement) warm-started from the previous psi, then ONE fused kernel that
// evaluates the whole tendency and the Shu-Osher stage combination
//   q_out = a q_n + b q_s + c dt rhs(q_s, psi)
// from an LDS tile (q with a 2-cell halo, psi with 1): HBM traffic per cell is
// read q_s, psi, q_n + write q_out.

#include <cmath>
#include <cstring>
#include <vector>

#include "fvx_common.hpp"
#include "mg.hpp"

namespace fvx {
template <typename T>
int unet_apply(void* net, const T* r, int64_t ld_r, T* z, int64_t ld_z, int H, int W,
               double out_scale, const uint8_t* mask, hipStream_t s, const int* guard);
template <typename T>
int lincomb3(int64_t n, double a, const T* x, double b, const T* y, double c, const T* z,
             T* out, hipStream_t s);
template <typename T>
int lincomb4(int64_t n, double a, const T* x, double b, const T* y, double c, const T* z,
             double d, const T* w, T* out, hipStream_t s);
}  // namespace fvx

extern "C" int fvx_mg_plan_create(const double*, const double*, int, int, double, double, double,
                                  int, int, int, int, double, int, int, void**);
extern "C" int fvx_mg_plan_destroy(void*);

namespace fvx {

constexpr int kQTx = 64, kQTy = 16, kQThreads = 256;

struct QGStageParams {
  double idx, idy, inv12, beta, dy, y0, F0, kF, nu4, a, b, cdt;
};

template <typename TS, typename TP, bool MASKED>
__global__ __launch_bounds__(kQThreads) void qg_stage_kernel(
    const TS* __restrict__ qs, const TP* __restrict__ psi, const TS* __restrict__ qn,
    TS* __restrict__ out, const uint8_t* __restrict__ mask, int Ny, int Nx, QGStageParams p,
    int tiles_x, int ntiles) {
  using T = TS;
  __shared__ T sq[kQTy + 4][kQTx + 4];
  __shared__ T sp[kQTy + 2][kQTx + 2];
  __shared__ T sl[kQTy + 2][kQTx + 2];
  __shared__ T sby[kQTy + 2];
  __shared__ T sF[kQTy];
  const int t = xcd_swizzle(blockIdx.x, ntiles);
  const int i0 = (t % tiles_x) * kQTx, j0 = (t / tiles_x) * kQTy;
  const int tid = threadIdx.x;
  const T idx = (T)p.idx, idy = (T)p.idy;

  for (int k = tid; k < (kQTy + 4) * (kQTx + 4); k += kQThreads) {
    const int yy = k / (kQTx + 4), xx = k % (kQTx + 4);
    const int gj = j0 - 2 + yy, gi = i0 - 2 + xx;
    sq[yy][xx] = (gj >= 0 && gj < Ny && gi >= 0 && gi < Nx) ? qs[(int64_t)gj * Nx + gi] : T(0);
  }
  for (int k = tid; k < (kQTy + 2) * (kQTx + 2); k += kQThreads) {
    const int yy = k / (kQTx + 2), xx = k % (kQTx + 2);
    const int gj = j0 - 1 + yy, gi = i0 - 1 + xx;
    sp[yy][xx] =
        (gj >= 0 && gj < Ny && gi >= 0 && gi < Nx) ? (T)psi[(int64_t)gj * Nx + gi] : T(0);
  }
  if (tid < kQTy + 2) {
    // beta (y - y0) at cell centres y_j = (j - 1/2) dy, rows j0-1 .. j0+TY.
    const int gj = j0 - 1 + tid;
    sby[tid] = (T)(p.beta * ((gj - 0.5) * p.dy - p.y0));
  } else if (tid >= 64 && tid < 64 + kQTy) {
    const int gj = j0 + tid - 64;
    sF[tid - 64] = (T)(-p.F0 * sin(p.kF * ((gj - 0.5) * p.dy)));
  }
  __syncthreads();

  // First no-flux Laplacian on the tile + 1 halo (zero outside the interior).
  for (int k = tid; k < (kQTy + 2) * (kQTx + 2); k += kQThreads) {
    const int yy = k / (kQTx + 2), xx = k % (kQTx + 2);
    const int gj = j0 - 1 + yy, gi = i0 - 1 + xx;
    T v = T(0);
    if (gj >= 1 && gj <= Ny - 2 && gi >= 1 && gi <= Nx - 2) {
      const int y = yy + 1, x = xx + 1;  // sq coordinates
      const T c = sq[y][x];
      const T fe = (gi <= Nx - 3) ? (sq[y][x + 1] - c) * idx : T(0);
      const T fw = (gi >= 2) ? (c - sq[y][x - 1]) * idx : T(0);
      const T fn = (gj <= Ny - 3) ? (sq[y + 1][x] - c) * idy : T(0);
      const T fs = (gj >= 2) ? (c - sq[y - 1][x]) * idy : T(0);
      v = (fe - fw) * idx + (fn - fs) * idy;
    }
    sl[yy][xx] = v;
  }
  __syncthreads();

  const T nu4 = (T)p.nu4, inv12 = (T)p.inv12;
  const T a = (T)p.a, b = (T)p.b, cdt = (T)p.cdt;
  for (int k = tid; k < kQTx * kQTy; k += kQThreads) {
    const int ty = k / kQTx, tx = k % kQTx;
    const int gj = j0 + ty, gi = i0 + tx;
    if (gj >= Ny || gi >= Nx) continue;
    const int64_t g = (int64_t)gj * Nx + gi;
    T res = T(0);
    if (gj >= 1 && gj <= Ny - 2 && gi >= 1 && gi <= Nx - 2) {
      // Arakawa J(psi, q + beta (y - y0)).
      const int y = ty + 1, x = tx + 1;  // sp / sl coordinates
      const int qy = ty + 2, qx = tx + 2;
      const T bN = sby[y + 1], bC = sby[y], bS = sby[y - 1];
      const T fE = sp[y][x + 1], fW = sp[y][x - 1], fN = sp[y + 1][x], fS = sp[y - 1][x];
      const T fNE = sp[y + 1][x + 1], fNW = sp[y + 1][x - 1], fSE = sp[y - 1][x + 1],
              fSW = sp[y - 1][x - 1];
      const T gE = sq[qy][qx + 1] + bC, gW = sq[qy][qx - 1] + bC;
      const T gN = sq[qy + 1][qx] + bN, gS = sq[qy - 1][qx] + bS;
      const T gNE = sq[qy + 1][qx + 1] + bN, gNW = sq[qy + 1][qx - 1] + bN;
      const T gSE = sq[qy - 1][qx + 1] + bS, gSW = sq[qy - 1][qx - 1] + bS;
      const T jpp = (fE - fW) * (gN - gS) - (fN - fS) * (gE - gW);
      const T jpx = fE * (gNE - gSE) - fW * (gNW - gSW) - fN * (gNE - gNW) + fS * (gSE - gSW);
      const T jxp = gN * (fNE - fNW) - gS * (fSE - fSW) - gE * (fNE - fSE) + gW * (fNW - fSW);
      const T J = ((jpp + jpx) + jxp) * inv12;
      // Second no-flux Laplacian -> biharmonic tendency -nu4 D(D q).
      const T c = sl[y][x];
      const T fe = (gi <= Nx - 3) ? (sl[y][x + 1] - c) * idx : T(0);
      const T fw = (gi >= 2) ? (c - sl[y][x - 1]) * idx : T(0);
      const T fn = (gj <= Ny - 3) ? (sl[y + 1][x] - c) * idy : T(0);
      const T fs = (gj >= 2) ? (c - sl[y - 1][x]) * idy : T(0);
      const T bih = -nu4 * ((fe - fw) * idx + (fn - fs) * idy);
      T rhs = (-J + sF[ty]) + bih;
      if constexpr (MASKED) rhs = rhs * (T)(mask[g] ? 1 : 0);
      const T qsv = sq[qy][qx];
      res = (a != T(0) ? a * qn[g] : T(0)) + b * qsv + cdt * rhs;
    }
    out[g] = res;
  }
}

// The same stage laid out for the wavefront: one wavefront per tile of 64
// columns (60 written, two halo lanes on each side) x kQWRows rows; each lane
// holds its column of q (rows j0-2 .. j0+R+1) and psi (rows j0-1 .. j0+R) in
// registers, east/west neighbours come from DPP lane shifts, north/south from
// the neighbouring registers: no LDS, no barriers.  Arithmetic as in
// qg_stage_kernel.  TS is the state's storage type, TC the type the tendency
// is computed in; the stage combination a q_n + b q_s + c dt rhs is formed in
// TS from the stored q_s, so an FP64 state with an FP32 tendency rounds the
// tendency only, never the state.
constexpr int kQWRows = 8;

template <typename T>
__device__ __forceinline__ T lane_e(T v) {
  return dpp_move<0x130>(v);  // wave_shl 1: value of lane + 1, 0 in lane 63
}
template <typename T>
__device__ __forceinline__ T lane_w(T v) {
  return dpp_move<0x138>(v);  // wave_shr 1: value of lane - 1, 0 in lane 0
}
template <typename T>
__device__ __forceinline__ T read_lane(T v, int l) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}
// Element (row_off + col) of a state array of n elements; FP32 arrays through
// a buffer descriptor (scalar row offset), FP64 through a plain load.
__device__ __forceinline__ float load_state(const float* a, int n, unsigned col, int row_off) {
  return buf_load(buf_rsrc(a, n), col, row_off);
}
__device__ __forceinline__ double load_state(const double* a, int n, unsigned col, int row_off) {
  (void)n;
  return a[(int64_t)row_off + col];
}

// Optional certification residual of the Helmholtz solve that produced psi,
// fused into the stage that consumes it: on every interior cell the FP64
// r = (q_s m - (lap(psi m) - lam psi m) m) m of the solve's operator (the
// refinement residual of mg.hip; psi is the solve's solution, q_s its
// right-hand side), and block partials of ||q_s m||^2 and ||r||^2
// (value-major, as the solver's finishing kernel reads them).
struct StageCert {
  double* partials;  // 2 x nblk doubles
  double lam, idx2, idy2;
};

// Block sums of (bb, rr) of a 256-thread block into partials[k * nblk + bid].
__device__ __forceinline__ void stage_partials(double bb, double rr, double* partials, int nblk,
                                               int bid) {
  __shared__ double red[4][2];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double s0 = wave_sum_last(bb), s1 = wave_sum_last(rr);
  if (lane == 63) {
    red[wid][0] = s0;
    red[wid][1] = s1;
  }
  __syncthreads();
  if (tid < 2) partials[(int64_t)tid * nblk + bid] = ((red[0][tid] + red[1][tid]) + red[2][tid]) +
                                                     red[3][tid];
}

// Body over block ``bx`` of ``nblk`` (four tiles per block); with MASKED a
// null mask means all wet.  With CERT the block also forms the certification
// residual partials of ``cert`` (block index bx of nblk).
template <typename TS, typename TC, typename TP, bool MASKED, bool CERT = false>
__device__ __forceinline__ void qg_stage_wave_body(
    const TS* __restrict__ qs, const TP* __restrict__ psi, const TS* __restrict__ qn,
    TS* __restrict__ out, const uint8_t* __restrict__ mask, int Ny, int Nx, const QGStageParams& p,
    int tiles_x, int ntiles, int bx, int nblk, const StageCert& cert = StageCert()) {
  using T = TC;
  constexpr int R = kQWRows;
  const int lane = threadIdx.x & 63;
  const int tile =
      __builtin_amdgcn_readfirstlane(xcd_swizzle(bx, nblk) * 4 + (threadIdx.x >> 6));
  const bool active = tile < ntiles;
  const int tyi = active ? tile / tiles_x : 0;
  const int i0 = (active ? tile - tyi * tiles_x : 0) * 60;
  const int j0 = tyi * R;
  const int gi = i0 - 2 + lane;
  const bool colin = active && gi >= 0 && gi < Nx;
  const unsigned gic = (unsigned)min(max(gi, 0), Nx - 1);
  const int n = Ny * Nx;
  // q rows j0-2+t (t = 0 .. R+3), psi rows j0-1+t (t = 0 .. R+1); zero
  // outside the array.  qc keeps the stored q_s of the R output rows.
  T q[R + 4], f[R + 2];
  TS qc[R];
#pragma unroll
  for (int t = 0; t < R + 4; ++t) {
    const int gj = j0 - 2 + t;
    TS v = load_state(qs, n, gic, min(max(gj, 0), Ny - 1) * Nx);
    v = (colin && gj >= 0 && gj < Ny) ? v : TS(0);
    q[t] = (T)v;
    if (t >= 2 && t < R + 2) qc[t - 2] = v;
  }
  // MASKED: bit t = the mask of row j0 - 1 + t in this lane's column (one
  // byte load per row, read once for the certification and the output).
  uint32_t mbits = ~0u;
  if constexpr (MASKED) {
    if (mask) {
      mbits = 0;
#pragma unroll
      for (int t = 0; t < R + 2; ++t) {
        const int gjc = min(max(j0 - 1 + t, 0), Ny - 1);
        mbits |= (mask[gjc * Nx + (int)gic] ? 1u : 0u) << t;
      }
    }
  }
  auto wet_t = [&](int t) -> bool { return MASKED ? ((mbits >> t) & 1u) != 0 : true; };
  // With CERT: psi * mask in FP64 on the interior, zero elsewhere (the
  // solve's masked, zero-padded operand); the row masks stay as bits.
  double xm[R + 2];
#pragma unroll
  for (int t = 0; t < R + 2; ++t) {
    const int gj = j0 - 1 + t;
    const int64_t g = (int64_t)min(max(gj, 0), Ny - 1) * Nx + gic;
    const TP v = psi[g];
    f[t] = (colin && gj >= 0 && gj < Ny) ? (T)v : T(0);
    if constexpr (CERT) {
      const bool in = colin && gj >= 1 && gj <= Ny - 2 && gi >= 1 && gi <= Nx - 2;
      xm[t] = (in && wet_t(t)) ? (double)v : 0.0;
    }
  }
  // beta (y - y0) at cell centres y_j = (j - 1/2) dy for rows j0-1 .. j0+R and
  // F = -F0 sin(kF y) for rows j0 .. j0+R-1: one row per lane, then read back
  // as wave-uniform values.
  const int gjl = j0 - 1 + lane;
  const T by_l = (T)(p.beta * ((gjl - 0.5) * p.dy - p.y0));
  const T F_l = (lane >= 1 && lane <= R) ? (T)(-p.F0 * sin(p.kF * ((gjl - 0.5) * p.dy))) : T(0);
  const T idx = (T)p.idx, idy = (T)p.idy;
  const bool ci = gi >= 1 && gi <= Nx - 2;  // interior column
  const bool cE = gi <= Nx - 3, cW = gi >= 2;
  // First no-flux Laplacian on rows j0-1 .. j0+R (zero outside the interior);
  // q's east/west neighbours kept for the Jacobian.
  T l1[R + 2], qE[R + 2], qW[R + 2];
#pragma unroll
  for (int t = 0; t < R + 2; ++t) {
    const int qi = t + 1, gj = j0 - 1 + t;
    const T c = q[qi];
    const T e = lane_e(c), w = lane_w(c);
    qE[t] = e;
    qW[t] = w;
    const T fe = cE ? (e - c) * idx : T(0);
    const T fw = cW ? (c - w) * idx : T(0);
    const T fn = (gj <= Ny - 3) ? (q[qi + 1] - c) * idy : T(0);
    const T fs = (gj >= 2) ? (c - q[qi - 1]) * idy : T(0);
    const bool in = ci && gj >= 1 && gj <= Ny - 2;
    l1[t] = in ? (fe - fw) * idx + (fn - fs) * idy : T(0);
  }
  T fE[R + 2], fW[R + 2];
#pragma unroll
  for (int t = 0; t < R + 2; ++t) {
    fE[t] = lane_e(f[t]);
    fW[t] = lane_w(f[t]);
  }
  const T nu4 = (T)p.nu4, inv12 = (T)p.inv12;
  const TS a = (TS)p.a, b = (TS)p.b, cdt = (TS)p.cdt;
  const bool lout = colin && lane >= 2 && lane < 62;
  double bb = 0.0, rr = 0.0;
#pragma unroll
  for (int t = 0; t < R; ++t) {
    if constexpr (CERT) {
      const int y = t + 1, gj = j0 + t;
      const double c = xm[y];
      const double e = lane_e(c), w = lane_w(c);
      if (lout && ci && gj >= 1 && gj <= Ny - 2) {
        const double lap = (e + w - 2.0 * c) * cert.idx2 + (xm[y + 1] + xm[y - 1] - 2.0 * c) * cert.idy2;
        const double m = wet_t(y) ? 1.0 : 0.0;
        const double bv = (double)qc[t] * m;
        const double rv = (bv - (lap - cert.lam * c) * m) * m;
        bb += bv * bv;
        rr += rv * rv;
      }
    }
    const int y = t + 1;  // f, l1, qE, qW, by index of this row
    const int qy = t + 2;  // q index of this row
    const int gj = j0 + t;
    // Arakawa J(psi, q + beta (y - y0)).
    const T bN = read_lane(by_l, y + 1), bC = read_lane(by_l, y), bS = read_lane(by_l, y - 1);
    const T fe_ = fE[y], fw_ = fW[y], fn_ = f[y + 1], fs_ = f[y - 1];
    const T fNE = fE[y + 1], fNW = fW[y + 1], fSE = fE[y - 1], fSW = fW[y - 1];
    const T gE = qE[y] + bC, gW = qW[y] + bC;
    const T gN = q[qy + 1] + bN, gS = q[qy - 1] + bS;
    const T gNE = qE[y + 1] + bN, gNW = qW[y + 1] + bN;
    const T gSE = qE[y - 1] + bS, gSW = qW[y - 1] + bS;
    const T jpp = (fe_ - fw_) * (gN - gS) - (fn_ - fs_) * (gE - gW);
    const T jpx = fe_ * (gNE - gSE) - fw_ * (gNW - gSW) - fn_ * (gNE - gNW) + fs_ * (gSE - gSW);
    const T jxp = gN * (fNE - fNW) - gS * (fSE - fSW) - gE * (fNE - fSE) + gW * (fNW - fSW);
    const T J = ((jpp + jpx) + jxp) * inv12;
    // Second no-flux Laplacian -> biharmonic tendency -nu4 D(D q).
    const T c = l1[y];
    const T le = lane_e(c), lw = lane_w(c);
    const T fe = cE ? (le - c) * idx : T(0);
    const T fw = cW ? (c - lw) * idx : T(0);
    const T fn = (gj <= Ny - 3) ? (l1[y + 1] - c) * idy : T(0);
    const T fs = (gj >= 2) ? (c - l1[y - 1]) * idy : T(0);
    const T bih = -nu4 * ((fe - fw) * idx + (fn - fs) * idy);
    const T rhs = (-J + read_lane(F_l, y)) + bih;
    if (lout && gj < Ny) {
      const int64_t g = (int64_t)gj * Nx + gi;
      TS res = TS(0);
      if (ci && gj >= 1 && gj <= Ny - 2) {
        TS r = (TS)rhs;
        if constexpr (MASKED) r = wet_t(t + 1) ? r : TS(0);
        res = (a != TS(0) ? a * qn[g] : TS(0)) + b * qc[t] + cdt * r;
      }
      out[g] = res;
    }
  }
  if constexpr (CERT) stage_partials(bb, rr, cert.partials, nblk, bx);
}

template <typename TS, typename TC, typename TP, bool MASKED, bool CERT = false>
__global__ __launch_bounds__(256) void qg_stage_wave_kernel(
    const TS* __restrict__ qs, const TP* __restrict__ psi, const TS* __restrict__ qn,
    TS* __restrict__ out, const uint8_t* __restrict__ mask, int Ny, int Nx, QGStageParams p,
    int tiles_x, int ntiles, StageCert cert) {
  qg_stage_wave_body<TS, TC, TP, MASKED, CERT>(qs, psi, qn, out, mask, Ny, Nx, p, tiles_x, ntiles,
                                               blockIdx.x, gridDim.x, cert);
}

// Work-list launch of the stage over many fields (ensemble members, layers):
// entry e owns blocks first[e] .. first[e+1]-1.
struct StageEntry {
  const double* qs;
  const double* psi;
  const double* qn;
  double* out;
  const uint8_t* mask;
  int Ny, Nx, tiles_x, ntiles;
  QGStageParams p;
};

__device__ __forceinline__ int stage_entry(const int* __restrict__ first, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename TC>
__global__ __launch_bounds__(256) void qg_stage_batch_kernel(const StageEntry* __restrict__ list,
                                                             const int* __restrict__ first,
                                                             int n) {
  const int e = stage_entry(first, n, blockIdx.x);
  const StageEntry& a = list[e];
  qg_stage_wave_body<double, TC, double, true>(a.qs, a.psi, a.qn, a.out, a.mask, a.Ny, a.Nx, a.p,
                                               a.tiles_x, a.ntiles, blockIdx.x - first[e],
                                               first[e + 1] - first[e]);
}

// Blocks of the wavefront-tile stage kernel on an Ny x Nx array.
static int stage_blocks(int Ny, int Nx) { return cdiv(cdiv(Nx, 60) * cdiv(Ny, kQWRows), 4); }

template <typename TS, typename TC, typename TP, bool MASKED>
static void launch_stage_wave(const TS* qs, const TP* psi, const TS* qn, TS* out,
                              const uint8_t* mask, int Ny, int Nx, const QGStageParams& p,
                              const StageCert* cert, hipStream_t s) {
  const int wtx = cdiv(Nx, 60), ntiles = wtx * cdiv(Ny, kQWRows);
  const dim3 grid(stage_blocks(Ny, Nx)), block(256);
  if (cert)
    hipLaunchKernelGGL((qg_stage_wave_kernel<TS, TC, TP, MASKED, true>), grid, block, 0, s, qs,
                       psi, qn, out, mask, Ny, Nx, p, wtx, ntiles, *cert);
  else
    hipLaunchKernelGGL((qg_stage_wave_kernel<TS, TC, TP, MASKED, false>), grid, block, 0, s, qs,
                       psi, qn, out, mask, Ny, Nx, p, wtx, ntiles, StageCert());
}

// tend32: compute the tendency of an FP64 state in FP32 (FP32 states always
// use FP32).  cert (FP64 states only): also form the certification residual
// partials of the solve that produced psi (stage_blocks(Ny, Nx) blocks).
template <typename TS, typename TP>
int qg_stage(const TS* qs, const TP* psi, const TS* qn, TS* out, const uint8_t* mask, int Ny,
             int Nx, const QGStageParams& p, hipStream_t s, bool tend32 = false,
             const StageCert* cert = nullptr) {
  FVX_CHECK_ARG(qs && psi && qn && out, "null field pointer");
  FVX_CHECK_ARG(Ny >= 3 && Nx >= 3, "array too small");
  FVX_CHECK_ARG(out != qs && out != (const TS*)psi, "out must not alias q_s or psi");
  FVX_CHECK_ARG(!cert || (std::is_same<TS, double>::value &&
                          (int64_t)Ny * Nx < (int64_t(1) << 29)),
                "the fused certification residual needs an FP64 state");
  if ((int64_t)Ny * Nx < (int64_t(1) << 29)) {
    if (std::is_same<TS, float>::value || tend32) {
      if (mask) launch_stage_wave<TS, float, TP, true>(qs, psi, qn, out, mask, Ny, Nx, p, cert, s);
      else launch_stage_wave<TS, float, TP, false>(qs, psi, qn, out, mask, Ny, Nx, p, cert, s);
    } else if constexpr (std::is_same<TS, double>::value) {
      if (mask) launch_stage_wave<TS, double, TP, true>(qs, psi, qn, out, mask, Ny, Nx, p, cert, s);
      else launch_stage_wave<TS, double, TP, false>(qs, psi, qn, out, mask, Ny, Nx, p, cert, s);
    }
    FVX_LAUNCH_CHECK();
    return kOk;
  }
  const int tx = cdiv(Nx, kQTx), ty = cdiv(Ny, kQTy);
  if (mask)
    hipLaunchKernelGGL((qg_stage_kernel<TS, TP, true>), dim3(tx * ty), dim3(kQThreads), 0, s, qs,
                       psi, qn, out, mask, Ny, Nx, p, tx, tx * ty);
  else
    hipLaunchKernelGGL((qg_stage_kernel<TS, TP, false>), dim3(tx * ty), dim3(kQThreads), 0, s, qs,
                       psi, qn, out, mask, Ny, Nx, p, tx, tx * ty);
  FVX_LAUNCH_CHECK();
  return kOk;
}

// ---------------------------------------------------------------------------
// QG plan and step driver
// ---------------------------------------------------------------------------

enum QGSolve : int { kSolveF64 = 64, kSolveMixed = 3264 };

struct QGPlan {
  int ny, nx, Ny, Nx;
  double Lx, Ly, dx, dy, beta, lam, nu4, F0, dt, rtol, inner_rtol;
  int state;  // 32 or 64
  bool tend32 = false;  // FP64 state with the tendency computed in FP32
  int solve;  // kSolveF64 or kSolveMixed
  int max_iters, max_outer;
  uint8_t* mask = nullptr;  // full-array wet mask (device), null = basin
  Plan* mg = nullptr;
  void* q1 = nullptr;
  void* q2 = nullptr;
  // Graph replay: ping-pong state buffers, each graph = one full step.
  void* qa = nullptr;
  void* qb = nullptr;
  // Streamfunctions of the three stages of the last step and of the one
  // before (four arrays B0..B3, rotating with period two): stage k of a step
  // solves into its own array, warm-started from an extrapolation of the
  // previous stages (stage_guess).
  double* ps[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int parity = 0;          // 0: B0 holds the last step's stage-0 psi, 1: B3
  bool fresh = true;       // no step taken since the history was reset
  const double* last_psi = nullptr;  // caller's psi written by the last call
  int cap_parity = 0;      // parity graph 0 was captured with
  int warm = 1;            // 1: extrapolated warm starts, 0: previous stage's psi
  hipStream_t cap = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  static constexpr int kMaxOuter = 8;
  // Iteration budgets of one solve in the step graphs: FP64 PCG iterations;
  // mixed: refinement steps and the inner iterations of each.
  struct Budget {
    int iters = 0, outer = 0;
    int inner[kMaxOuter] = {};
  };
  Budget kb[3];            // per RK stage, as captured
  Budget obs[3];           // per stage: largest counts seen (eager steps)
  Budget win[3];           // per stage: largest counts of the current window
  int b_inner = 0;         // inner budget of refinement steps never observed
  int win_steps = 0;       // converged graph steps in the current window
  int win_len = 4;         // window length before budgets are tightened
  bool calibrated = false;
  int eager_steps = 0;  // synchronous steps run before calibration
  int recaptures = 0;
  // Learned preconditioner for the inner FP32 solves (mixed mode only).
  void* net = nullptr;
  double net_scale = 1.0;
  int net_sweeps = 0;
  int fallbacks = 0;
};

// PrecondFn adapter: z = net(r) on the interior grid.
static int net_precond(void* ctx, const float* r, float* z, int ny, int nx, const int* guard,
                       hipStream_t s) {
  (void)ny;
  (void)nx;
  QGPlan* Q = (QGPlan*)ctx;
  return learned_precond_apply(*Q->mg, Q->net, Q->net_scale, Q->net_sweeps, r, z, guard, s);
}
struct QGInfoC {
  int solves, iterations, max_iterations, outer;
  int converged, retries;
  double max_rel_residual;
};

static QGStageParams stage_params(const QGPlan& Q, double a, double b, double c) {
  QGStageParams p;
  p.idx = 1.0 / Q.dx;
  p.idy = 1.0 / Q.dy;
  p.inv12 = 1.0 / (12.0 * Q.dx * Q.dy);
  p.beta = Q.beta;
  p.dy = Q.dy;
  p.y0 = 0.5 * Q.Ly;
  p.F0 = Q.F0;
  p.kF = 2.0 * M_PI / Q.Ly;
  p.nu4 = Q.nu4;
  p.a = a;
  p.b = b;
  p.cdt = c * Q.dt;
  return p;
}

// Warm start of stage k's Helmholtz solve in a step taken with parity e,
// formed in the array that stage solves into.  psi is linear in q, so the
// Shu-Osher stages of rk3_ssp_step give the streamfunction of each stage
// state from earlier solutions and the streamfunctions of dt times the
// tendency, D(y) = psi(dt L(y)).  With P, N the stage-0 solutions of the
// previous and this step and B1, B2 the arrays of stages 1 and 2 (holding the
// previous step's until overwritten), the previous step gives exactly
//   D0 = D(q_{n-1}) = B1 - P            (tau = 0, in units of dt from t_{n-1})
//   D1 = D(y1)      = 4 B2 - 3 P - B1   (tau ~ 1)
//   D2 = D(y2)      = (3 N - P - 2 B2) / 2   (tau ~ 1/2, once N is known)
// and each stage needs one unknown D, taken linear in tau:
//   stage 0: N  = P/3 + 2 B2/3 + 2/3 D(y2),  D(y2) ~ (D0 + D1)/2  ->  2 B2 - P
//   stage 1: B1 = N + D(q_n),  D(q_n) ~ 2 D2 - D0              ->  4 N - 2 B2 - B1
//   stage 2: B2 = 3N/4 + B1/4 + D(y1_n)/4 (B1 now this step's), D(y1_n) from
//            D(q_n) = B1 - N and D2 at tau 1 and 1/2 ->  B1 - 3N/4 + P/4 + B2/2
// The three guesses leave initial residuals 3-5x smaller than taking psi
// linear in time over the stage times (measured on the 1.5-layer basin).
// The first step after a reset has no P: stage 0 starts from the last psi,
// stage 1 from N, stage 2 from B1.  warm == 0 starts every stage from the
// previous stage's solution.  Returns the stage's streamfunction array.
// The six history arrays in the roles of a step of parity e: the stage-0
// solutions of the previous and this step (P, N), and the stage-1 / stage-2
// solutions of the previous step (B1o, B2o) and of this one (B1n, B2n); the
// roles alternate with the parity, so no warm start overwrites an array it
// reads.
struct StageBufs {
  double *P, *N, *B1o, *B1n, *B2o, *B2n;
};
static StageBufs stage_bufs(const QGPlan& Q, int e) {
  StageBufs b;
  b.P = Q.ps[e == 0 ? 0 : 3];
  b.N = Q.ps[e == 0 ? 3 : 0];
  b.B1o = Q.ps[e == 0 ? 1 : 4];
  b.B1n = Q.ps[e == 0 ? 4 : 1];
  b.B2o = Q.ps[e == 0 ? 2 : 5];
  b.B2n = Q.ps[e == 0 ? 5 : 2];
  return b;
}

struct GuessTerms {
  const double* h[4];
  double c[4];
  int n;
};

// Terms of stage k's warm start (x = sum g.c[t] g.h[t]) and the array the
// stage solves into.
static double* stage_guess_terms(const QGPlan& Q, int k, int e, GuessTerms* g) {
  const StageBufs b = stage_bufs(Q, e);
  const bool extrap = Q.warm == 1 && !Q.fresh;
  auto set = [&](int t, double c, const double* h) {
    g->c[t] = c;
    g->h[t] = h;
  };
  for (int t = 0; t < 4; ++t) set(t, 0.0, nullptr);
  if (k == 0) {
    if (extrap) {
      set(0, 2.0, b.B2o);
      set(1, -1.0, b.P);
      g->n = 2;
    } else {
      set(0, 1.0, b.B2o);
      g->n = 1;
    }
    return b.N;
  }
  if (k == 1) {
    if (extrap) {
      set(0, 4.0, b.N);
      set(1, -2.0, b.B2o);
      set(2, -1.0, b.B1o);
      g->n = 3;
    } else {
      set(0, 1.0, b.N);
      g->n = 1;
    }
    return b.B1n;
  }
  if (extrap) {
    set(0, 1.0, b.B1n);
    set(1, -0.75, b.N);
    set(2, 0.25, b.P);
    set(3, 0.5, b.B2o);
    g->n = 4;
  } else {
    set(0, 1.0, b.B1n);
    g->n = 1;
  }
  return b.B2n;
}

static double* stage_guess(QGPlan& Q, int k, int e, hipStream_t s, int* rc) {
  const int64_t n = (int64_t)Q.Ny * Q.Nx;
  GuessTerms g;
  double* x = stage_guess_terms(Q, k, e, &g);
  *rc = lincomb4<double>(n, g.c[0], g.h[0], g.c[1], g.h[1], g.c[2], g.h[2], g.c[3], g.h[3], x, s);
  return x;
}

// Take the caller's psi as the warm start unless it is the array the last
// call returned (then the plan's stage history continues).
static int sync_history(QGPlan& Q, const double* psi, hipStream_t s) {
  if (psi == Q.last_psi && !Q.fresh) return kOk;
  const size_t bytes = (size_t)Q.Ny * Q.Nx * sizeof(double);
  for (int b = 0; b < 6; ++b)
    FVX_HIP(hipMemcpyAsync(Q.ps[b], psi, bytes, hipMemcpyDeviceToDevice, s));
  Q.fresh = true;
  return kOk;
}

// The last stage's streamfunction: B2n of the last step (B2o of the next).
static int return_psi(QGPlan& Q, double* psi, hipStream_t s) {
  FVX_HIP(hipMemcpyAsync(psi, stage_bufs(Q, Q.parity).B2o, (size_t)Q.Ny * Q.Nx * sizeof(double),
                         hipMemcpyDeviceToDevice, s));
  Q.last_psi = psi;
  return kOk;
}

template <typename TS>
static int qg_step_impl(QGPlan& Q, TS* q, int nsteps, QGInfoC* info, hipStream_t s) {
  const int Nx = Q.Nx;
  const int64_t off = (int64_t)Nx + 1;  // interior origin
  TS* q1 = (TS*)Q.q1;
  TS* q2 = (TS*)Q.q2;
  const double coef[3][3] = {{0.0, 1.0, 1.0}, {0.75, 0.25, 0.25},
                             {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
  QGInfoC acc;
  std::memset(&acc, 0, sizeof(acc));
  acc.converged = 1;
  for (int st = 0; st < nsteps; ++st) {
    for (int k = 0; k < 3; ++k) {
      const TS* qs = (k == 0) ? q : (k == 1 ? q1 : q2);
      TS* out = (k == 0) ? q1 : (k == 1 ? q2 : q);
      CgInfoC ci;
      std::memset(&ci, 0, sizeof(ci));
      int rc = kOk;
      double* psi = stage_guess(Q, k, Q.parity, s, &rc);
      if (rc) return rc;
      if (Q.solve == kSolveMixed) {
        rc = mixed_solve<TS>(*Q.mg, qs + off, Nx, psi + off, Nx, Q.rtol, 0.0, Q.max_outer,
                             Q.max_iters, Q.inner_rtol, &ci, 0, s,
                             Q.net ? &net_precond : nullptr, Q.net ? &Q : nullptr);
        Q.fallbacks += ci.pad;
        if (!rc) {
          QGPlan::Budget& ob = Q.obs[k];
          Q.b_inner = std::max(Q.b_inner, plan_max_inner(*Q.mg, 0));
          int per[QGPlan::kMaxOuter];
          const int no = plan_inner_iters(*Q.mg, 0, per, QGPlan::kMaxOuter);
          for (int o = 0; o < no; ++o) ob.inner[o] = std::max(ob.inner[o], per[o]);
          ob.outer = std::max(ob.outer, ci.outer);
        }
      } else {
        if constexpr (std::is_same<TS, double>::value) {
          // Recursive residual to rtol/2 so the certified FP64 true residual
          // meets rtol despite recurrence drift; certified at rtol, as the
          // graph path certifies.
          rc = pcg_solve<double>(*Q.mg, qs + off, Nx, psi + off, Nx, 0.5 * Q.rtol, 0.0,
                                 Q.max_iters, true, &ci, 0, s, Q.rtol);
        } else {
          set_error("FP64 MG-PCG needs an FP64 state");
          return kInvalidArgument;
        }
      }
      if (rc) return rc;
      Q.obs[k].iters = std::max(Q.obs[k].iters, ci.iterations);
      acc.solves += 1;
      acc.iterations += ci.iterations;
      acc.max_iterations = std::max(acc.max_iterations, ci.iterations);
      acc.outer += ci.outer;
      if (!ci.converged) acc.converged = 0;
      const double rel = ci.rhs_norm > 0 ? ci.residual_norm / ci.rhs_norm : ci.residual_norm;
      acc.max_rel_residual = std::max(acc.max_rel_residual, rel);
      rc = qg_stage<TS, double>(qs, psi, q, out, Q.mask, Q.Ny, Q.Nx,
                                stage_params(Q, coef[k][0], coef[k][1], coef[k][2]), s,
                                Q.tend32);
      if (rc) return rc;
    }
    Q.parity ^= 1;
    Q.fresh = false;
  }
  if (info) *info = acc;
  return kOk;
}

// Enqueue one whole step q_in -> q_out with parity e (all kernels, no host
// work): the sequence a graph captures.  Solves use the plan's budgets and
// slots 0..2.
template <typename TS>
static int qg_step_enqueue(QGPlan& Q, const TS* q_in, TS* q_out, int e, hipStream_t s) {
  const int Nx = Q.Nx;
  const int64_t off = (int64_t)Nx + 1;
  TS* q1 = (TS*)Q.q1;
  TS* q2 = (TS*)Q.q2;
  const double coef[3][3] = {{0.0, 1.0, 1.0}, {0.75, 0.25, 0.25},
                             {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
  for (int k = 0; k < 3; ++k) {
    const TS* qs = (k == 0) ? q_in : (k == 1 ? q1 : q2);
    TS* out = (k == 0) ? q1 : (k == 1 ? q2 : q_out);
    int rc = kOk;
    double* psi = stage_guess(Q, k, e, s, &rc);
    if (rc) return rc;
    // FP64 states: the solve's closing FP64 residual is formed inside the
    // stage kernel that reads the same q_s and psi (one pass less per stage);
    // its finishing kernel then certifies the solve as before.  The stage's
    // output is only kept when the host finds every solve certified.
    const bool fuse = std::is_same<TS, double>::value;
    StageCert cert;
    if (fuse) {
      size_t cap = 0;
      cert.partials = plan_partials(*Q.mg, &cap);
      cert.lam = Q.lam;
      cert.idx2 = 1.0 / (Q.dx * Q.dx);
      cert.idy2 = 1.0 / (Q.dy * Q.dy);
      FVX_CHECK_ARG(cap >= (size_t)2 * stage_blocks(Q.Ny, Q.Nx), "reduction scratch too small");
    }
    if (Q.solve == kSolveMixed) {
      rc = mixed_enqueue<TS>(*Q.mg, qs + off, Nx, psi + off, Nx, Q.rtol, 0.0, Q.kb[k].outer,
                             Q.kb[k].inner, Q.inner_rtol, k, s, Q.net ? &net_precond : nullptr,
                             Q.net ? &Q : nullptr, fuse);
    } else {
      if constexpr (std::is_same<TS, double>::value) {
        rc = pcg_enqueue<double>(*Q.mg, qs + off, Nx, psi + off, Nx, 0.5 * Q.rtol, 0.0, true,
                                 Q.kb[k].iters, k, s);
      } else {
        set_error("FP64 MG-PCG needs an FP64 state");
        return kInvalidArgument;
      }
    }
    if (rc) return rc;
    rc = qg_stage<TS, double>(qs, psi, q_in, out, Q.mask, Q.Ny, Q.Nx,
                              stage_params(Q, coef[k][0], coef[k][1], coef[k][2]), s, Q.tend32,
                              fuse ? &cert : nullptr);
    if (rc) return rc;
    if (fuse)
      certify_finish(*Q.mg, stage_blocks(Q.Ny, Q.Nx), Q.rtol, 0.0, k, Q.solve != kSolveMixed, s);
  }
  return kOk;
}

static void drop_graphs(QGPlan& Q) {
  for (int g = 0; g < 2; ++g) {
    if (Q.exec[g]) (void)hipGraphExecDestroy(Q.exec[g]);
    if (Q.graph[g]) (void)hipGraphDestroy(Q.graph[g]);
    Q.exec[g] = nullptr;
    Q.graph[g] = nullptr;
  }
}

template <typename TS>
static int capture_graphs(QGPlan& Q) {
  drop_graphs(Q);
  for (int g = 0; g < 2; ++g) {
    const TS* in = (const TS*)(g == 0 ? Q.qa : Q.qb);
    TS* out = (TS*)(g == 0 ? Q.qb : Q.qa);
    FVX_HIP(hipStreamBeginCapture(Q.cap, hipStreamCaptureModeThreadLocal));
    int rc = qg_step_enqueue<TS>(Q, in, out, Q.cap_parity ^ g, Q.cap);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(Q.cap, &gr);
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) {
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return kHipError;
    }
    Q.graph[g] = gr;
    FVX_HIP(hipGraphInstantiate(&Q.exec[g], gr, nullptr, nullptr, 0));
  }
  return kOk;
}

// Convergence of the three solves of the last replayed step, and the counts
// each stage's solve used (per refinement step for the mixed solve): a
// refinement step's inner count is recorded when the next refinement
// residual resets the inner solve; the last budgeted step's stays in iter.
static bool step_converged(QGPlan& Q, QGInfoC& acc, hipStream_t s, bool stage_ok[3],
                           QGPlan::Budget used[3]) {
  const PcgScalars* h = fetch_slots(*Q.mg, 0, 3, s);
  if (!h) return false;
  bool ok = true;
  for (int k = 0; k < 3; ++k) {
    const PcgScalars& in = h[2 * k];
    const PcgScalars& out = h[2 * k + 1];
    const bool conv = out.out_done != 0;  // FP64 true residual (both modes)
    stage_ok[k] = conv;
    ok = ok && conv;
    QGPlan::Budget& u = used[k];
    u = QGPlan::Budget();
    if (Q.solve == kSolveMixed) {
      u.outer = out.outer;
      for (int o = 0; o < std::min(out.outer, QGPlan::kMaxOuter); ++o) {
        if (o == Q.kb[k].outer - 1) u.inner[o] = in.iter;
        else u.inner[o] = (o < 4) ? in.it_outer[o] : Q.kb[k].inner[o];
      }
    } else {
      u.iters = in.iter;
    }
    acc.solves += 1;
    acc.iterations += in.total;
    acc.max_iterations = std::max(acc.max_iterations, in.total);
    acc.outer += out.outer;
    const double rel = out.bb > 0 ? std::sqrt(out.out_rr / out.bb) : std::sqrt(out.out_rr);
    acc.max_rel_residual = std::max(acc.max_rel_residual, rel);
  }
  return ok;
}

static void budget_max(QGPlan::Budget& a, const QGPlan::Budget& b) {
  a.iters = std::max(a.iters, b.iters);
  a.outer = std::max(a.outer, b.outer);
  for (int o = 0; o < QGPlan::kMaxOuter; ++o) a.inner[o] = std::max(a.inner[o], b.inner[o]);
}

// Largest budgets over the stages (reported by fvx_qg_budgets).
static QGPlan::Budget budget_all(const QGPlan& Q) {
  QGPlan::Budget m;
  for (int k = 0; k < 3; ++k) budget_max(m, Q.kb[k]);
  return m;
}

template <typename TS>
static int qg_run_graph(QGPlan& Q, TS* q, double* psi, int nsteps, QGInfoC* info,
                        hipStream_t s) {
  const size_t N = (size_t)Q.Ny * Q.Nx;
  QGInfoC acc;
  std::memset(&acc, 0, sizeof(acc));
  acc.converged = 1;
  if (nsteps == 0) {
    if (info) *info = acc;
    return kOk;
  }
  if (int rc = sync_history(Q, psi, s)) return rc;
  int done_steps = 0;
  TS* cur = q;
  auto absorb = [&](const QGInfoC& ci) {
    acc.solves += ci.solves;
    acc.iterations += ci.iterations;
    acc.max_iterations = std::max(acc.max_iterations, ci.max_iterations);
    acc.outer += ci.outer;
    acc.max_rel_residual = std::max(acc.max_rel_residual, ci.max_rel_residual);
    if (!ci.converged) acc.converged = 0;
  };
  const int outer_cap = std::min(Q.max_outer, QGPlan::kMaxOuter);
  // Calibration: synchronous steps observe the iteration counts, then the
  // step graphs are captured with those budgets (also within a one-step
  // call, so that later calls replay graphs only).  The first steps usually
  // start from a cold psi (zero or stale) and need more iterations than the
  // run will, so the first kObserveStep - 1 steps run unobserved; the next
  // one is observed (mixed precision: with unchunked inner solves, so each
  // refinement step shows the iterations it needs from a warm start).
  constexpr int kObserveStep = 3;
  while (!Q.calibrated && done_steps < nsteps) {
    const bool observe = Q.eager_steps >= kObserveStep - 1;
    int so = 0, si = 0;
    if (observe) {
      for (int k = 0; k < 3; ++k) Q.obs[k] = QGPlan::Budget();
      Q.b_inner = 0;
      if (Q.solve == kSolveMixed) {
        plan_schedule(*Q.mg, &so, &si);
        plan_set_schedule(*Q.mg, 1, Q.max_iters);
      }
    }
    QGInfoC ci;
    int rc = qg_step_impl<TS>(Q, q, 1, &ci, s);
    if (observe && Q.solve == kSolveMixed) plan_set_schedule(*Q.mg, so, si);
    if (rc) return rc;
    absorb(ci);
    ++done_steps;
    ++Q.eager_steps;
    if (!observe) continue;
    // Budgets: each stage's observed counts.  Multigrid-preconditioned solves
    // take about the same count step after step, so they get no margin:
    // every iteration slot a converged solve leaves unused still costs its
    // launches (~30 us at 2048^2), and a step whose solve misses its budget
    // is redone with a wider one.  The learned preconditioner's longer, less
    // regular counts keep one iteration of margin.  Refinement steps never
    // observed get the largest inner count seen, plus one.
    const int margin = Q.net ? 1 : 0;
    Q.b_inner = std::max(2, Q.b_inner + 1);
    for (int k = 0; k < 3; ++k) {
      const QGPlan::Budget& ob = Q.obs[k];
      QGPlan::Budget& b = Q.kb[k];
      b.iters = std::max(2, ob.iters + margin);
      b.outer = std::min(std::max(1, ob.outer), QGPlan::kMaxOuter);
      for (int o = 0; o < QGPlan::kMaxOuter; ++o)
        b.inner[o] = ob.inner[o] > 0 ? std::min(Q.max_iters, ob.inner[o] + margin) : Q.b_inner;
      Q.win[k] = QGPlan::Budget();
    }
    Q.win_steps = 0;
    Q.cap_parity = Q.parity;
    rc = capture_graphs<TS>(Q);
    if (rc) return rc;
    Q.calibrated = true;
  }
  // The step graphs extrapolate warm starts from the previous step's stage
  // solutions: after a reset of the history one step runs eagerly first.
  if (done_steps < nsteps && Q.fresh) {
    QGInfoC ci;
    int rc = qg_step_impl<TS>(Q, q, 1, &ci, s);
    if (rc) return rc;
    absorb(ci);
    ++done_steps;
  }
  if (done_steps < nsteps) {
    if (Q.exec[0] == nullptr) {
      Q.cap_parity = Q.parity;
      int rc = capture_graphs<TS>(Q);
      if (rc) return rc;
    }
    // exec[0]: qa -> qb, exec[1]: qb -> qa; graph g was captured for steps
    // of parity cap_parity ^ g.
    int g = Q.parity ^ Q.cap_parity;
    FVX_HIP(hipMemcpyAsync(g == 0 ? Q.qa : Q.qb, q, N * sizeof(TS), hipMemcpyDeviceToDevice, s));
    while (done_steps < nsteps) {
      FVX_HIP(hipGraphLaunch(Q.exec[g], s));
      QGInfoC st;
      std::memset(&st, 0, sizeof(st));
      bool stage_ok[3];
      QGPlan::Budget used[3];
      if (step_converged(Q, st, s, stage_ok, used)) {
        acc.solves += st.solves;
        acc.iterations += st.iterations;
        acc.max_iterations = std::max(acc.max_iterations, st.max_iterations);
        acc.outer += st.outer;
        acc.max_rel_residual = std::max(acc.max_rel_residual, st.max_rel_residual);
        ++done_steps;
        g ^= 1;
        Q.parity ^= 1;
        // Tighten: after win_len converged steps, each stage's budgets come
        // down to the largest counts its solves used in that window (the
        // learned preconditioner keeps its margin).
        for (int k = 0; k < 3; ++k) budget_max(Q.win[k], used[k]);
        if (++Q.win_steps >= Q.win_len && !Q.net) {
          bool tighter = false;
          for (int k = 0; k < 3; ++k) {
            QGPlan::Budget& b = Q.kb[k];
            const QGPlan::Budget& w = Q.win[k];
            if (Q.solve == kSolveMixed) {
              const int outer = std::max(1, w.outer);
              if (outer < b.outer) {
                b.outer = outer;
                tighter = true;
              }
              for (int o = 0; o < b.outer; ++o) {
                const int inner = std::max(1, w.inner[o]);
                if (inner < b.inner[o]) {
                  b.inner[o] = inner;
                  tighter = true;
                }
              }
            } else if (std::max(1, w.iters) < b.iters) {
              b.iters = std::max(1, w.iters);
              tighter = true;
            }
            Q.win[k] = QGPlan::Budget();
          }
          Q.win_steps = 0;
          if (tighter) {
            Q.cap_parity = Q.parity ^ g;
            int rc = capture_graphs<TS>(Q);
            if (rc) return rc;
          }
        }
        continue;
      }
      // A solve missed its tolerance within the budget: widen that stage's
      // budgets and redo the step from its (untouched) input buffer; the
      // next tightening waits twice as long.
      const QGPlan::Budget top = budget_all(Q);
      int top_inner = 0;
      for (int o = 0; o < top.outer; ++o) top_inner = std::max(top_inner, top.inner[o]);
      if (Q.net && (top_inner >= Q.max_iters || top.outer >= Q.max_outer)) {
        // The learned preconditioner cannot meet the tolerance within the
        // caps: fall back to multigrid-PCG for the rest of the run.
        Q.net = nullptr;
        Q.fallbacks += 1;
        acc.retries += 1;
        Q.cap_parity = Q.parity ^ g;
        int rc = capture_graphs<TS>(Q);
        if (rc) return rc;
        continue;
      }
      // Give up widening once the budget that can still grow is at its cap
      // (FP64: PCG iterations; mixed: refinement steps and inner iterations).
      bool capped = true;
      for (int k = 0; k < 3; ++k) {
        if (stage_ok[k]) continue;
        const QGPlan::Budget& b = Q.kb[k];
        bool inner_capped = true;
        for (int o = 0; o < b.outer; ++o) inner_capped = inner_capped && b.inner[o] >= Q.max_iters;
        capped = capped && ((Q.solve == kSolveMixed) ? (b.outer >= outer_cap && inner_capped)
                                                      : (b.iters >= Q.max_iters));
      }
      if (Q.recaptures > 50 || capped) {
        acc.converged = 0;
        acc.solves += st.solves;
        acc.max_rel_residual = std::max(acc.max_rel_residual, st.max_rel_residual);
        ++done_steps;
        g ^= 1;
        Q.parity ^= 1;
        continue;
      }
      Q.recaptures += 1;
      acc.retries += 1;
      for (int k = 0; k < 3; ++k) {
        if (stage_ok[k]) continue;
        QGPlan::Budget& b = Q.kb[k];
        b.iters = std::min(Q.max_iters, b.iters + 2);
        // widen by a quarter (at least one iteration) so long budgets
        // converge in a few recaptures
        for (int o = 0; o < QGPlan::kMaxOuter; ++o)
          b.inner[o] = std::min(Q.max_iters, b.inner[o] + std::max(1, b.inner[o] / 4));
        b.outer = std::min(outer_cap, b.outer + 1);
      }
      for (int k = 0; k < 3; ++k) Q.win[k] = QGPlan::Budget();
      Q.win_steps = 0;
      Q.win_len = std::min(64, 2 * Q.win_len);
      Q.cap_parity = Q.parity ^ g;
      int rc = capture_graphs<TS>(Q);
      if (rc) return rc;
    }
    cur = (TS*)(g == 0 ? Q.qa : Q.qb);
    FVX_HIP(hipMemcpyAsync(q, cur, N * sizeof(TS), hipMemcpyDeviceToDevice, s));
  }
  if (int rc = return_psi(Q, psi, s)) return rc;
  FVX_HIP(hipStreamSynchronize(s));
  if (info) *info = acc;
  return kOk;
}

// ---------------------------------------------------------------------------
// Geostrophic face velocities of a T-point streamfunction (the composition
// of scripts/qg_1p5_layer.py:115-145): psi is averaged to X points
// (Interpolation2D.T_to_X, interior only, times mask.xy_corner_strict), then
// u = -diff_y_X_to_U(psi_X) (times mask.u) and v = diff_x_X_to_V(psi_X)
// (times mask.v), both on the interior with a zero ring.  One thread per
// cell computes the three corner values it needs from nine psi loads.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void geostrophic_faces_kernel(
    const double* __restrict__ psi, const uint8_t* __restrict__ mxy,
    const uint8_t* __restrict__ mu, const uint8_t* __restrict__ mv, T* __restrict__ u,
    T* __restrict__ v, int ny, int nx, double dx, double dy) {
  const int64_t n = (int64_t)ny * nx;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += stride) {
    const int j = (int)(c / nx), i = (int)(c - (int64_t)j * nx);
    auto X = [&](int jj, int ii) -> double {
      if (jj < 1 || jj > ny - 2 || ii < 1 || ii > nx - 2) return 0.0;
      const int64_t k = (int64_t)jj * nx + ii;
      double x = 0.25 * (psi[k] + psi[k + 1] + psi[k + nx] + psi[k + nx + 1]);
      if (mxy) x = x * (double)(mxy[k] ? 1 : 0);
      return x;
    };
    double uu = 0.0, vv = 0.0;
    if (j >= 1 && j <= ny - 2 && i >= 1 && i <= nx - 2) {
      const double xc = X(j, i);
      uu = -((xc - X(j - 1, i)) / dy);
      vv = (xc - X(j, i - 1)) / dx;
      if (mu) uu = uu * (double)(mu[c] ? 1 : 0);
      if (mv) vv = vv * (double)(mv[c] ? 1 : 0);
    }
    u[c] = (T)uu;
    v[c] = (T)vv;
  }
}

template <typename T>
static int geostrophic_faces(const double* psi, const uint8_t* mxy, const uint8_t* mu,
                             const uint8_t* mv, T* u, T* v, int ny, int nx, double dx, double dy,
                             hipStream_t s) {
  FVX_CHECK_ARG(psi && u && v, "null field pointer");
  FVX_CHECK_ARG((void*)u != (void*)psi && (void*)v != (void*)psi && (void*)u != (void*)v,
                "u, v and psi must be distinct arrays");
  FVX_CHECK_ARG(ny >= 3 && nx >= 3, "array too small (%d x %d)", ny, nx);
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  FVX_CHECK_ARG((mxy == nullptr) == (mu == nullptr) && (mu == nullptr) == (mv == nullptr),
                "mask_xy, mask_u and mask_v go together");
  const int64_t n = (int64_t)ny * nx;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(geostrophic_faces_kernel<T>, dim3(blocks), dim3(256), 0, s, psi, mxy, mu,
                     mv, u, v, ny, nx, dx, dy);
  FVX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace fvx

extern "C" {

typedef struct fvx_qg_info {
  int solves;           // Helmholtz solves performed
  int iterations;       // total inner PCG iterations
  int max_iterations;   // largest per-solve inner iteration count
  int outer;            // total refinement sweeps (mixed)
  int converged;        // 1 if every solve met the tolerance
  int retries;          // graph steps redone with a wider iteration budget
  double max_rel_residual;  // max over solves of ||b - A psi|| / ||b|| (FP64)
} fvx_qg_info;

int fvx_qg_plan_create(int ny, int nx, double Lx, double Ly, double beta, double Ld, double nu4,
                       double F0, double dt, const double* mask_interior, int state_precision,
                       int solve_mode, int n_pre, int n_post, int n_coarse, double omega,
                       double rtol, double inner_rtol, int max_iters, int max_outer,
                       void** plan_out) {
  using namespace fvx;
  FVX_CHECK_ARG(plan_out, "null plan_out");
  FVX_CHECK_ARG(ny >= 8 && nx >= 8, "interior must be at least 8 x 8");
  FVX_CHECK_ARG(state_precision == 32 || state_precision == 64 || state_precision == 6432,
                "state precision 32, 64 or 6432 (FP64 state, FP32 tendency)");
  FVX_CHECK_ARG(solve_mode == kSolveF64 || solve_mode == kSolveMixed,
                "solve mode must be 64 (FP64 MG-PCG) or 3264 (mixed)");
  FVX_CHECK_ARG(!(solve_mode == kSolveF64 && state_precision != 64),
                "the FP64 MG-PCG path needs an FP64 state and FP64 tendency");
  FVX_CHECK_ARG(Lx > 0 && Ly > 0 && Ld > 0 && dt > 0, "physical parameters must be positive");
  QGPlan* Q = new QGPlan();
  Q->ny = ny;
  Q->nx = nx;
  Q->Ny = ny + 2;
  Q->Nx = nx + 2;
  Q->Lx = Lx;
  Q->Ly = Ly;
  Q->dx = Lx / nx;
  Q->dy = Ly / ny;
  Q->beta = beta;
  Q->lam = 1.0 / (Ld * Ld);
  Q->nu4 = nu4;
  Q->F0 = F0;
  Q->dt = dt;
  Q->rtol = rtol;
  Q->inner_rtol = inner_rtol;
  Q->state = state_precision == 32 ? 32 : 64;
  Q->tend32 = state_precision == 6432;
  Q->solve = solve_mode;
  Q->max_iters = max_iters;
  Q->max_outer = max_outer;
  std::vector<double> m((size_t)ny * nx, 1.0);
  bool masked = false;
  if (mask_interior) {
    for (size_t k = 0; k < m.size(); ++k) {
      m[k] = mask_interior[k] != 0.0 ? 1.0 : 0.0;
      if (m[k] == 0.0) masked = true;
    }
  }
  void* mg = nullptr;
  // Hierarchy depth: coarsen until the cell size reaches Ld / 2 (lambda h^2
  // >= 1/4 on the bottom level, which gets n_coarse damped-Jacobi sweeps).
  // With the extrapolated warm starts the PCG iteration count is the same as
  // for a hierarchy down to 2 Ld (1.73 per solve at 2048^2, Ld = 16 cells)
  // while each V-cycle has two fewer levels of latency-bound launches; one
  // level fewer again (Ld / 4) needs more iterations and misses budgets.
  int n_levels = 0;  // 0 = as deep as the grid halves
  if (Q->lam > 0.0) {
    const double h = std::min(Q->dx, Q->dy);
    const double ratio = 0.5 / (std::sqrt(Q->lam) * h);  // Ld / (2 h)
    n_levels = std::max(2, 1 + (int)std::ceil(std::log2(std::max(ratio, 1.0))));
    int deepest = 1, hy = ny, hx = nx;  // the halving rule of multigrid.py:50-76
    while (hy % 2 == 0 && hx % 2 == 0 && hy / 2 >= 8 && hx / 2 >= 8) {
      hy /= 2;
      hx /= 2;
      ++deepest;
    }
    n_levels = std::min(n_levels, deepest);
  }
  int rc = fvx_mg_plan_create(m.data(), nullptr, ny, nx, Q->dx, Q->dy, Q->lam, n_levels, n_pre,
                              n_post, n_coarse, omega, 1 /*dirichlet*/,
                              solve_mode == kSolveMixed ? 3264 : 64, &mg);
  if (rc) {
    delete Q;
    return rc;
  }
  Q->mg = (Plan*)mg;
  // Chebyshev weights for the inner FP32 V-cycles' sweeps where the
  // hierarchy takes them: on [0.25, 2] with land, [0.4, 2] without.
  // Measured at 2048^2 (3 sweeps a leg, Ld = 16 cells): 1407 -> 1465 steps/s
  // all wet, 847 -> 983 on the basin (2.91 -> 2.51 inner iterations a solve);
  // lo = 0.1 (the best for cold-start solves) costs iterations against the
  // extrapolated warm starts.
  if (solve_mode == kSolveMixed && plan_set_chebyshev(*Q->mg, masked ? 0.25 : 0.4) != kOk)
    (void)plan_set_chebyshev(*Q->mg, 0.0);
  const size_t N = (size_t)Q->Ny * Q->Nx;
  const size_t es = Q->state == 64 ? 8 : 4;
  hipError_t e = hipMalloc(&Q->q1, N * es);
  if (e == hipSuccess) e = hipMalloc(&Q->q2, N * es);
  if (e == hipSuccess) e = hipMalloc(&Q->qa, N * es);
  if (e == hipSuccess) e = hipMalloc(&Q->qb, N * es);
  for (int b = 0; b < 6 && e == hipSuccess; ++b) {
    e = hipMalloc(&Q->ps[b], N * sizeof(double));
    if (e == hipSuccess) e = hipMemset(Q->ps[b], 0, N * sizeof(double));
  }
  if (e == hipSuccess) e = hipMemset(Q->qa, 0, N * es);
  if (e == hipSuccess) e = hipMemset(Q->qb, 0, N * es);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&Q->cap, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemset(Q->q1, 0, N * es);
  if (e == hipSuccess) e = hipMemset(Q->q2, 0, N * es);
  if (e == hipSuccess && masked) {
    std::vector<uint8_t> mf(N, 0);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) mf[(size_t)(j + 1) * Q->Nx + (i + 1)] = m[(size_t)j * nx + i] != 0.0;
    e = hipMalloc(&Q->mask, N);
    if (e == hipSuccess) e = hipMemcpy(Q->mask, mf.data(), N, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    set_error("HIP allocation failed: %s", hipGetErrorString(e));
    fvx_mg_plan_destroy(mg);
    delete Q;
    return kHipError;
  }
  *plan_out = Q;
  return kOk;
}

int fvx_qg_plan_destroy(void* plan) {
  using namespace fvx;
  if (!plan) return kOk;
  QGPlan* Q = (QGPlan*)plan;
  (void)hipDeviceSynchronize();
  drop_graphs(*Q);
  if (Q->cap) (void)hipStreamDestroy(Q->cap);
  fvx_mg_plan_destroy(Q->mg);
  (void)hipFree(Q->q1);
  (void)hipFree(Q->q2);
  (void)hipFree(Q->qa);
  (void)hipFree(Q->qb);
  for (int b = 0; b < 6; ++b) (void)hipFree(Q->ps[b]);
  (void)hipFree(Q->mask);
  delete Q;
  return kOk;
}

// Advance q (full (ny+2) x (nx+2) array of the plan's state precision, ghost
// ring zero) by nsteps rk3_ssp_step steps; psi (FP64, same shape) is the warm
// start on entry and the streamfunction of the last stage on exit.
int fvx_qg_step(void* plan, void* q, double* psi, int nsteps, fvx_qg_info* info, void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(plan && q && psi, "null argument");
  FVX_CHECK_ARG(nsteps >= 0, "negative nsteps");
  QGPlan* Q = (QGPlan*)plan;
  hipStream_t s = (hipStream_t)stream;
  if (int rc = sync_history(*Q, psi, s)) return rc;
  int rc = (Q->state == 64)
               ? qg_step_impl<double>(*Q, (double*)q, nsteps, (QGInfoC*)info, s)
               : qg_step_impl<float>(*Q, (float*)q, nsteps, (QGInfoC*)info, s);
  if (rc) return rc;
  return return_psi(*Q, psi, s);
}

// Warm start of each stage's Helmholtz solve: 1 (default) extrapolates from
// the previous stages' solutions, 0 starts from the previous stage's
// solution.  Invalidates captured graphs.
int fvx_qg_set_warm_start(void* plan, int mode) {
  using namespace fvx;
  FVX_CHECK_ARG(plan, "null plan");
  FVX_CHECK_ARG(mode == 0 || mode == 1, "warm-start mode must be 0 or 1");
  QGPlan* Q = (QGPlan*)plan;
  Q->warm = mode;
  drop_graphs(*Q);
  Q->calibrated = false;
  Q->eager_steps = 0;
  return kOk;
}

// Coarse-level wall correction of the inner FP32 V-cycles of a
// mixed-precision plan (fvx_mg_plan_set_wall); on = 1 enables it where every
// level runs on the tiled kernels and returns 1 if it is on, 0 if the plan's
// hierarchy cannot take it (or on = 0).  Invalidates captured graphs.
int fvx_qg_set_wall(void* plan, int on, int* enabled) {
  using namespace fvx;
  FVX_CHECK_ARG(plan && enabled, "null argument");
  QGPlan* Q = (QGPlan*)plan;
  *enabled = 0;
  if (Q->solve == kSolveMixed) {
    if (on && plan_set_wall(*Q->mg, true) == kOk) *enabled = 1;
    else (void)plan_set_wall(*Q->mg, false);
  }
  drop_graphs(*Q);
  Q->calibrated = false;
  Q->eager_steps = 0;
  return kOk;
}

// Chebyshev smoothing weights (fvx_mg_plan_set_chebyshev) for the V-cycles
// of the plan's solves (the inner FP32 V-cycles of a mixed-precision plan,
// the FP64 V-cycles of an FP64 one); *enabled = 1 when on.  Invalidates
// captured graphs.
int fvx_qg_set_chebyshev(void* plan, double lo, int* enabled) {
  using namespace fvx;
  FVX_CHECK_ARG(plan && enabled, "null argument");
  QGPlan* Q = (QGPlan*)plan;
  *enabled = 0;
  if (lo > 0.0 && plan_set_chebyshev(*Q->mg, lo) == kOk) *enabled = 1;
  else (void)plan_set_chebyshev(*Q->mg, 0.0);
  drop_graphs(*Q);
  Q->calibrated = false;
  Q->eager_steps = 0;
  return kOk;
}

// Use a learned U-Net (fvx_unet_create) as the preconditioner of the inner
// FP32 solves of a mixed-precision plan (net = NULL restores multigrid);
// z = out_scale * rms(r) * net(r / rms(r)).  out_scale <= 0 selects the
// inverse Jacobi diagonal of the fine operator.  Invalidates captured graphs.
int fvx_qg_set_preconditioner(void* plan, void* net, double out_scale, int sweeps) {
  using namespace fvx;
  FVX_CHECK_ARG(plan, "null plan");
  QGPlan* Q = (QGPlan*)plan;
  FVX_CHECK_ARG(net == nullptr || Q->solve == kSolveMixed,
                "a learned preconditioner needs the mixed-precision solve");
  FVX_CHECK_ARG(sweeps >= 0 && sweeps <= 8, "sweeps must be in [0, 8]");
  if (net) {  // allocate the work arrays outside any graph capture
    int nyc = 0, nxc = 0;
    FVX_CHECK_ARG(learned_coarse_dims(*Q->mg, &nyc, &nxc) == 0,
                  "the learned preconditioner needs at least two multigrid levels");
    for (int k = 0; k < 4; ++k)
      FVX_CHECK_ARG(fine_work(*Q->mg, k) != nullptr, "work array allocation failed");
  }
  Q->net = net;
  Q->net_sweeps = sweeps;
  Q->net_scale = out_scale > 0 ? out_scale : plan_inv_diag0(*Q->mg);
  drop_graphs(*Q);
  Q->calibrated = false;
  Q->fallbacks = 0;
  return kOk;
}

int fvx_qg_fallbacks(void* plan) { return plan ? ((fvx::QGPlan*)plan)->fallbacks : -1; }

// Iteration budgets of the captured step graphs: out[0] FP64 PCG iterations,
// out[1] inner FP32 iterations per refinement sweep, out[2] refinement
// sweeps, out[3] recaptures so far.
int fvx_qg_budgets(void* plan, int* out) {
  using namespace fvx;
  FVX_CHECK_ARG(plan && out, "null argument");
  QGPlan* Q = (QGPlan*)plan;
  const QGPlan::Budget b = budget_all(*Q);
  int inner = 0;
  for (int o = 0; o < b.outer; ++o) inner = std::max(inner, b.inner[o]);
  out[0] = b.iters;
  out[1] = inner;
  out[2] = b.outer;
  out[3] = Q->recaptures;
  return kOk;
}

// Same as fvx_qg_step, replaying one captured HIP graph per step: each graph
// holds the three solves (fixed iteration budgets, device-side convergence
// guards, FP64 certification) and the three stage kernels.  The host checks
// every solve's FP64 residual flag after each step and redoes a step with a
// wider budget when one is missed.  The first call calibrates the budgets
// with one synchronous step.
int fvx_qg_step_graph(void* plan, void* q, double* psi, int nsteps, fvx_qg_info* info,
                      void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(plan && q && psi, "null argument");
  FVX_CHECK_ARG(nsteps >= 0, "negative nsteps");
  QGPlan* Q = (QGPlan*)plan;
  hipStream_t s = (hipStream_t)stream;
  if (Q->state == 64)
    return qg_run_graph<double>(*Q, (double*)q, psi, nsteps, (QGInfoC*)info, s);
  return qg_run_graph<float>(*Q, (float*)q, psi, nsteps, (QGInfoC*)info, s);
}

typedef struct fvx_qg_stage_desc {
  const double* qs;
  const double* psi;
  const double* qn;
  double* out;
  const uint8_t* mask;  // full-array wet mask or NULL
  int Ny, Nx;
  double dx, dy, beta, Ly, F0, nu4, a, b, cdt;
} fvx_qg_stage_desc;

struct StageBatchHandle {
  fvx::StageEntry* list = nullptr;
  int* first = nullptr;
  int n = 0, blocks = 0, tend32 = 0;
};

// A fixed work-list of QG stage evaluations (FP64 states): one launch runs
// every entry (ensemble members, layers of a multi-layer model).
int fvx_qg_stage_batch_create(int n, const fvx_qg_stage_desc* d, int tend32, void** out) {
  using namespace fvx;
  FVX_CHECK_ARG(n > 0 && d && out, "null argument or empty list");
  std::vector<StageEntry> e(n);
  std::vector<int> first(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    const fvx_qg_stage_desc& x = d[k];
    FVX_CHECK_ARG(x.qs && x.psi && x.qn && x.out, "entry %d: null field pointer", k);
    FVX_CHECK_ARG(x.out != x.qs && x.out != x.psi, "entry %d: out aliases q_s or psi", k);
    FVX_CHECK_ARG(x.Ny >= 3 && x.Nx >= 3 && (int64_t)x.Ny * x.Nx < (int64_t(1) << 29),
                  "entry %d: bad shape %d x %d", k, x.Ny, x.Nx);
    StageEntry& a = e[k];
    a.qs = x.qs;
    a.psi = x.psi;
    a.qn = x.qn;
    a.out = x.out;
    a.mask = x.mask;
    a.Ny = x.Ny;
    a.Nx = x.Nx;
    a.tiles_x = cdiv(x.Nx, 60);
    a.ntiles = a.tiles_x * cdiv(x.Ny, kQWRows);
    a.p.idx = 1.0 / x.dx;
    a.p.idy = 1.0 / x.dy;
    a.p.inv12 = 1.0 / (12.0 * x.dx * x.dy);
    a.p.beta = x.beta;
    a.p.dy = x.dy;
    a.p.y0 = 0.5 * x.Ly;
    a.p.F0 = x.F0;
    a.p.kF = 2.0 * M_PI / x.Ly;
    a.p.nu4 = x.nu4;
    a.p.a = x.a;
    a.p.b = x.b;
    a.p.cdt = x.cdt;
    first[k + 1] = first[k] + cdiv(a.ntiles, 4);
  }
  StageBatchHandle* H = new StageBatchHandle();
  H->n = n;
  H->blocks = first[n];
  H->tend32 = tend32 ? 1 : 0;
  hipError_t err = hipMalloc(&H->list, sizeof(StageEntry) * n);
  if (err == hipSuccess) err = hipMalloc(&H->first, sizeof(int) * (n + 1));
  if (err == hipSuccess)
    err = hipMemcpy(H->list, e.data(), sizeof(StageEntry) * n, hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = hipMemcpy(H->first, first.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(H->list);
    (void)hipFree(H->first);
    delete H;
    set_error("HIP error %s", hipGetErrorString(err));
    return kHipError;
  }
  *out = H;
  return kOk;
}

int fvx_qg_stage_batch_launch(void* h, void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(h, "null handle");
  StageBatchHandle* H = (StageBatchHandle*)h;
  if (H->tend32)
    hipLaunchKernelGGL(qg_stage_batch_kernel<float>, dim3(H->blocks), dim3(256), 0,
                       (hipStream_t)stream, H->list, H->first, H->n);
  else
    hipLaunchKernelGGL(qg_stage_batch_kernel<double>, dim3(H->blocks), dim3(256), 0,
                       (hipStream_t)stream, H->list, H->first, H->n);
  FVX_LAUNCH_CHECK();
  return kOk;
}

int fvx_qg_stage_batch_destroy(void* h) {
  if (!h) return 0;
  StageBatchHandle* H = (StageBatchHandle*)h;
  (void)hipDeviceSynchronize();
  (void)hipFree(H->list);
  (void)hipFree(H->first);
  delete H;
  return 0;
}

#define FVX_QG_STAGE(SUF, TS, T32)                                                                \
  int fvx_qg_stage_##SUF(const TS* qs, const double* psi, const TS* qn, TS* out,                  \
                         const uint8_t* mask, int Ny, int Nx, double dx, double dy, double beta,  \
                         double Ly, double F0, double nu4, double a, double b, double cdt,        \
                         void* stream) {                                                          \
    using namespace fvx;                                                                          \
    QGStageParams p;                                                                              \
    p.idx = 1.0 / dx;                                                                             \
    p.idy = 1.0 / dy;                                                                             \
    p.inv12 = 1.0 / (12.0 * dx * dy);                                                             \
    p.beta = beta;                                                                                \
    p.dy = dy;                                                                                    \
    p.y0 = 0.5 * Ly;                                                                              \
    p.F0 = F0;                                                                                    \
    p.kF = 2.0 * M_PI / Ly;                                                                       \
    p.nu4 = nu4;                                                                                  \
    p.a = a;                                                                                      \
    p.b = b;                                                                                      \
    p.cdt = cdt;                                                                                  \
    return qg_stage<TS, double>(qs, psi, qn, out, mask, Ny, Nx, p, (hipStream_t)stream, T32);     \
  }
FVX_QG_STAGE(f64, double, false)
FVX_QG_STAGE(f32, float, true)
// FP64 state, tendency computed in FP32, combination in FP64.
FVX_QG_STAGE(f64t32, double, true)
#undef FVX_QG_STAGE

// The same stage with the fused certification residual of the Helmholtz
// solve (lap - lam) psi = q_s on the interior: block partials of ||q_s m||^2
// (partials[0 .. nblk-1]) and ||q_s m - A psi||^2 (partials[nblk .. 2nblk-1]).
#define FVX_QG_STAGE_CERT(SUF, T32)                                                               \
  int fvx_qg_stage_cert_##SUF(const double* qs, const double* psi, const double* qn, double* out, \
                              const uint8_t* mask, int Ny, int Nx, double dx, double dy,         \
                              double beta, double Ly, double F0, double nu4, double a, double b, \
                              double cdt, double lam, double* partials, int64_t partials_cap,    \
                              int* nblk, void* stream) {                                         \
    using namespace fvx;                                                                          \
    FVX_CHECK_ARG(partials && nblk, "null partials or nblk");                                     \
    FVX_CHECK_ARG(Ny >= 3 && Nx >= 3 && dx > 0 && dy > 0, "bad grid");                           \
    const int nb = stage_blocks(Ny, Nx);                                                          \
    FVX_CHECK_ARG(partials_cap >= 2 * (int64_t)nb, "partials needs %d doubles", 2 * nb);         \
    QGStageParams p;                                                                              \
    p.idx = 1.0 / dx;                                                                             \
    p.idy = 1.0 / dy;                                                                             \
    p.inv12 = 1.0 / (12.0 * dx * dy);                                                             \
    p.beta = beta;                                                                                \
    p.dy = dy;                                                                                    \
    p.y0 = 0.5 * Ly;                                                                              \
    p.F0 = F0;                                                                                    \
    p.kF = 2.0 * M_PI / Ly;                                                                       \
    p.nu4 = nu4;                                                                                  \
    p.a = a;                                                                                      \
    p.b = b;                                                                                      \
    p.cdt = cdt;                                                                                  \
    StageCert c;                                                                                  \
    c.partials = partials;                                                                        \
    c.lam = lam;                                                                                  \
    c.idx2 = 1.0 / (dx * dx);                                                                     \
    c.idy2 = 1.0 / (dy * dy);                                                                     \
    *nblk = nb;                                                                                   \
    return qg_stage<double, double>(qs, psi, qn, out, mask, Ny, Nx, p, (hipStream_t)stream, T32,  \
                                    &c);                                                          \
  }
FVX_QG_STAGE_CERT(f64, false)
FVX_QG_STAGE_CERT(f64t32, true)
#undef FVX_QG_STAGE_CERT

int fvx_geostrophic_faces_f64(const double* psi, const uint8_t* mask_xy, const uint8_t* mask_u,
                              const uint8_t* mask_v, double* u, double* v, int ny, int nx,
                              double dx, double dy, void* stream) {
  return fvx::geostrophic_faces<double>(psi, mask_xy, mask_u, mask_v, u, v, ny, nx, dx, dy,
                                        (hipStream_t)stream);
}
int fvx_geostrophic_faces_f32(const double* psi, const uint8_t* mask_xy, const uint8_t* mask_u,
                              const uint8_t* mask_v, float* u, float* v, int ny, int nx,
                              double dx, double dy, void* stream) {
  return fvx::geostrophic_faces<float>(psi, mask_xy, mask_u, mask_v, u, v, ny, nx, dx, dy,
                                       (hipStream_t)stream);
}
}
