// C-grid stencils: Difference2D / Divergence2D / Interpolation2D point ops,
// Diffusion2D (+ fluxes), fused BiharmonicDiffusion2D, the Arakawa Jacobian,
// masked_laplacian and the periodic ghost fill.
//
// All kernels are HBM-bandwidth bound; each output element is written once
// and each input element is read from HBM once (neighbour reuse comes from
// L1/LDS).  Point ops use one thread per output on a 64 x 4 workgroup so a
// wavefront covers 64 contiguous elements of a row.

#include "fvx_common.hpp"

namespace fvx {

// Operator codes of the point-stencil family; the order is part of the C ABI
// (include/fvx.h, FVX_OP_*).
enum StencilOp : int {
  kDiffX_T2U = 0,   // difference.py:194  (h[j,i+1]-h[j,i])/dx        mask u
  kDiffY_T2V = 1,   // difference.py:215  (h[j+1,i]-h[j,i])/dy        mask v
  kDiffY_U2X = 2,   // difference.py:236                                mask X strict
  kDiffX_V2X = 3,   // difference.py:258                                mask X strict
  kDiffY_X2U = 4,   // difference.py:280  (q[j,i]-q[j-1,i])/dy         mask u
  kDiffX_X2V = 5,   // difference.py:301  (q[j,i]-q[j,i-1])/dx         mask v
  kDiffX_U2T = 6,   // difference.py:326                                mask h
  kDiffY_V2T = 7,   // difference.py:347                                mask h
  kDivergence = 8,  // difference.py:372 / _divergence_2d :62          mask h
  kDivNoflux = 9,   // divergence.py:138 (wall faces zeroed first)      mask h
  kCurl = 10,       // difference.py:401 / _curl_2d :47                 mask X strict
  kLaplacian = 11,  // difference.py:431                                mask h
  kGradPerpU = 12,  // difference.py:503-507                            mask u
  kGradPerpV = 13,  // difference.py:510-514                            mask v
  kT2X = 14,        // interpolation.py:158                             mask X strict
  kV2T = 15,        // interpolation.py:206                             mask h
  kNumOps = 16,
};

template <typename T, int OP>
__global__ __launch_bounds__(256) void point_stencil_kernel(
    const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out,
    const uint8_t* __restrict__ mask, int ny, int nx, T idx, T idy) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int64_t c = (int64_t)j * nx + i;
  T v = T(0);
  if (j >= 1 && j <= ny - 2 && i >= 1 && i <= nx - 2) {
    if constexpr (OP == kDiffX_T2U || OP == kDiffX_V2X) {
      v = (a[c + 1] - a[c]) * idx;
    } else if constexpr (OP == kDiffY_T2V || OP == kDiffY_U2X) {
      v = (a[c + nx] - a[c]) * idy;
    } else if constexpr (OP == kDiffY_X2U || OP == kDiffY_V2T) {
      v = (a[c] - a[c - nx]) * idy;
    } else if constexpr (OP == kDiffX_X2V || OP == kDiffX_U2T) {
      v = (a[c] - a[c - 1]) * idx;
    } else if constexpr (OP == kDivergence) {
      v = (a[c] - a[c - 1]) * idx + (b[c] - b[c - nx]) * idy;
    } else if constexpr (OP == kDivNoflux) {
      // u[:,0] = u[:,-2] = 0 and v[0,:] = v[-2,:] = 0 before the divergence.
      const T ue = (i == nx - 2) ? T(0) : a[c];
      const T uw = (i - 1 == 0) ? T(0) : a[c - 1];
      const T vn = (j == ny - 2) ? T(0) : b[c];
      const T vs = (j - 1 == 0) ? T(0) : b[c - nx];
      v = (ue - uw) * idx + (vn - vs) * idy;
    } else if constexpr (OP == kCurl) {
      v = (b[c + 1] - b[c]) * idx - (a[c + nx] - a[c]) * idy;
    } else if constexpr (OP == kLaplacian) {
      const T h0 = a[c];
      v = ((a[c + 1] - h0) - (h0 - a[c - 1])) * (idx * idx) +
          ((a[c + nx] - h0) - (h0 - a[c - nx])) * (idy * idy);
    } else if constexpr (OP == kGradPerpU) {
      v = -(a[c + nx] + a[c + nx + 1] - a[c - nx] - a[c - nx + 1]) * (T(0.25) * idy);
    } else if constexpr (OP == kGradPerpV) {
      v = (a[c + 1] + a[c + nx + 1] - a[c - 1] - a[c + nx - 1]) * (T(0.25) * idx);
    } else if constexpr (OP == kT2X) {
      v = T(0.25) * (a[c] + a[c + 1] + a[c + nx] + a[c + nx + 1]);
    } else if constexpr (OP == kV2T) {
      v = T(0.5) * (a[c] + a[c - nx]);
    }
  }
  if (mask) v = v * (T)(mask[c] ? 1 : 0);
  out[c] = v;
}

template <typename T>
using PointKernel = void (*)(const T*, const T*, T*, const uint8_t*, int, int, T, T);

template <typename T, int... OPS>
struct PointTable {
  static constexpr PointKernel<T> table[] = {&point_stencil_kernel<T, OPS>...};
};

template <typename T>
int point_stencil(int op, const T* a, const T* b, T* out, int ny, int nx, double dx,
                  double dy, const uint8_t* mask, hipStream_t s) {
  FVX_CHECK_ARG(op >= 0 && op < kNumOps, "unknown stencil op %d", op);
  FVX_CHECK_ARG(a && out, "null field pointer");
  FVX_CHECK_ARG(ny >= 3 && nx >= 3, "array too small (%d x %d); need >= 3 x 3", ny, nx);
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  const bool two = (op == kDivergence || op == kDivNoflux || op == kCurl);
  FVX_CHECK_ARG(!two || b, "operator needs a second field");
  using Tab = PointTable<T, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15>;
  dim3 blk(64, 4), grd(cdiv(nx, 64), cdiv(ny, 4));
  hipLaunchKernelGGL(Tab::table[op], grd, blk, 0, s, a, b, out, mask, ny, nx,
                     (T)(1.0 / dx), (T)(1.0 / dy));
  FVX_LAUNCH_CHECK();
  return kOk;
}

// ---------------------------------------------------------------------------
// Diffusion2D (diffusion.py:120-176): face fluxes exist only strictly between
// interior cells (no-flux walls); masked fluxes; tendency x mask.h.
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T kappa_at(const T* kap, T k0, int64_t c) {
  return kap ? kap[c] : k0;
}

template <typename T>
__device__ __forceinline__ T flux_x_at(const T* h, const T* kap, T k0, const uint8_t* mu,
                                       int j, int i, int ny, int nx, T idx) {
  // flux_x[j, i+1/2] for i = 1..nx-3, rows 1..ny-2 (diffusion.py:151-155)
  if (j < 1 || j > ny - 2 || i < 1 || i > nx - 3) return T(0);
  const int64_t c = (int64_t)j * nx + i;
  T f = kappa_at(kap, k0, c) * (h[c + 1] - h[c]) * idx;
  if (mu) f = f * (T)(mu[c] ? 1 : 0);
  return f;
}

template <typename T>
__device__ __forceinline__ T flux_y_at(const T* h, const T* kap, T k0, const uint8_t* mv,
                                       int j, int i, int ny, int nx, T idy) {
  // flux_y[j+1/2, i] for j = 1..ny-3, cols 1..nx-2 (diffusion.py:159-165)
  if (j < 1 || j > ny - 3 || i < 1 || i > nx - 2) return T(0);
  const int64_t c = (int64_t)j * nx + i;
  T f = kappa_at(kap, k0, c) * (h[c + nx] - h[c]) * idy;
  if (mv) f = f * (T)(mv[c] ? 1 : 0);
  return f;
}

template <typename T>
__global__ __launch_bounds__(256) void diffusion_kernel(
    const T* __restrict__ h, const T* __restrict__ kap, T k0, T* __restrict__ out,
    const uint8_t* __restrict__ mh, const uint8_t* __restrict__ mu,
    const uint8_t* __restrict__ mv, int ny, int nx, T idx, T idy) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int64_t c = (int64_t)j * nx + i;
  T v = T(0);
  if (j >= 1 && j <= ny - 2 && i >= 1 && i <= nx - 2) {
    const T fe = flux_x_at(h, kap, k0, mu, j, i, ny, nx, idx);
    const T fw = flux_x_at(h, kap, k0, mu, j, i - 1, ny, nx, idx);
    const T fn = flux_y_at(h, kap, k0, mv, j, i, ny, nx, idy);
    const T fs = flux_y_at(h, kap, k0, mv, j - 1, i, ny, nx, idy);
    v = (fe - fw) * idx + (fn - fs) * idy;
  }
  if (mh) v = v * (T)(mh[c] ? 1 : 0);
  out[c] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void diffusion_fluxes_kernel(
    const T* __restrict__ h, const T* __restrict__ kap, T k0, T* __restrict__ fx,
    T* __restrict__ fy, const uint8_t* __restrict__ mu, const uint8_t* __restrict__ mv,
    int ny, int nx, T idx, T idy) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int64_t c = (int64_t)j * nx + i;
  fx[c] = flux_x_at(h, kap, k0, mu, j, i, ny, nx, idx);
  fy[c] = flux_y_at(h, kap, k0, mv, j, i, ny, nx, idy);
}

template <typename T>
int diffusion(const T* h, const T* kap, double k0, T* out, int ny, int nx, double dx,
              double dy, const uint8_t* mh, const uint8_t* mu, const uint8_t* mv,
              hipStream_t s) {
  FVX_CHECK_ARG(h && out, "null field pointer");
  FVX_CHECK_ARG(ny >= 3 && nx >= 3, "array too small (%d x %d)", ny, nx);
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  dim3 blk(64, 4), grd(cdiv(nx, 64), cdiv(ny, 4));
  hipLaunchKernelGGL(diffusion_kernel<T>, grd, blk, 0, s, h, kap, (T)k0, out, mh, mu, mv, ny,
                     nx, (T)(1.0 / dx), (T)(1.0 / dy));
  FVX_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int diffusion_fluxes(const T* h, const T* kap, double k0, T* fx, T* fy, int ny, int nx,
                     double dx, double dy, const uint8_t* mu, const uint8_t* mv,
                     hipStream_t s) {
  FVX_CHECK_ARG(h && fx && fy, "null field pointer");
  FVX_CHECK_ARG(ny >= 3 && nx >= 3, "array too small (%d x %d)", ny, nx);
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  dim3 blk(64, 4), grd(cdiv(nx, 64), cdiv(ny, 4));
  hipLaunchKernelGGL(diffusion_fluxes_kernel<T>, grd, blk, 0, s, h, kap, (T)k0, fx, fy, mu,
                     mv, ny, nx, (T)(1.0 / dx), (T)(1.0 / dy));
  FVX_LAUNCH_CHECK();
  return kOk;
}

// ---------------------------------------------------------------------------
// Fused BiharmonicDiffusion2D (diffusion.py:543-574): -kappa * D(D(h)) with the
// unmasked no-flux Laplacian D applied twice through an LDS tile; the
// intermediate never touches HBM.  Traffic: read h, write out (2 x sizeof T
// per cell, + 1 byte when masked).
// ---------------------------------------------------------------------------

constexpr int kBhTx = 64, kBhTy = 8;  // output tile; 512 threads

// No-flux Laplacian of a field at (j, i) given a fetch functor; the fluxes
// between (j,i) and a neighbour exist only when both are interior cells.
template <typename T, typename F>
__device__ __forceinline__ T noflux_lap(F fetch, int j, int i, int ny, int nx, T idx,
                                        T idy) {
  if (j < 1 || j > ny - 2 || i < 1 || i > nx - 2) return T(0);
  const T c = fetch(j, i);
  const T fe = (i <= nx - 3) ? (fetch(j, i + 1) - c) * idx : T(0);
  const T fw = (i >= 2) ? (c - fetch(j, i - 1)) * idx : T(0);
  const T fn = (j <= ny - 3) ? (fetch(j + 1, i) - c) * idy : T(0);
  const T fs = (j >= 2) ? (c - fetch(j - 1, i)) * idy : T(0);
  return (fe - fw) * idx + (fn - fs) * idy;
}

template <typename T>
__global__ __launch_bounds__(512) void biharmonic_kernel(
    const T* __restrict__ h, T* __restrict__ out, const uint8_t* __restrict__ mh, int ny,
    int nx, T idx, T idy, T kappa, int tiles_x, int ntiles) {
  __shared__ T sh[kBhTy + 4][kBhTx + 4];
  __shared__ T sl[kBhTy + 2][kBhTx + 2];
  const int t = xcd_swizzle(blockIdx.x, ntiles);
  const int i0 = (t % tiles_x) * kBhTx, j0 = (t / tiles_x) * kBhTy;
  const int tid = threadIdx.x;
  for (int k = tid; k < (kBhTy + 4) * (kBhTx + 4); k += 512) {
    const int jj = k / (kBhTx + 4), ii = k % (kBhTx + 4);
    const int gj = j0 - 2 + jj, gi = i0 - 2 + ii;
    sh[jj][ii] = (gj >= 0 && gj < ny && gi >= 0 && gi < nx) ? h[(int64_t)gj * nx + gi] : T(0);
  }
  __syncthreads();
  auto fh = [&](int gj, int gi) { return sh[gj - j0 + 2][gi - i0 + 2]; };
  for (int k = tid; k < (kBhTy + 2) * (kBhTx + 2); k += 512) {
    const int jj = k / (kBhTx + 2), ii = k % (kBhTx + 2);
    sl[jj][ii] = noflux_lap<T>(fh, j0 - 1 + jj, i0 - 1 + ii, ny, nx, idx, idy);
  }
  __syncthreads();
  auto fl = [&](int gj, int gi) { return sl[gj - j0 + 1][gi - i0 + 1]; };
  const int tx = tid % kBhTx, ty = tid / kBhTx;
  const int gi = i0 + tx, gj = j0 + ty;
  if (gi < nx && gj < ny) {
    T v = -kappa * noflux_lap<T>(fl, gj, gi, ny, nx, idx, idy);
    const int64_t c = (int64_t)gj * nx + gi;
    if (mh) v = v * (T)(mh[c] ? 1 : 0);
    out[c] = v;
  }
}

template <typename T>
int biharmonic(const T* h, T* out, int ny, int nx, double dx, double dy, double kappa,
               const uint8_t* mh, hipStream_t s) {
  FVX_CHECK_ARG(h && out, "null field pointer");
  FVX_CHECK_ARG(ny >= 3 && nx >= 3, "array too small (%d x %d)", ny, nx);
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  const int tx = cdiv(nx, kBhTx), ty = cdiv(ny, kBhTy);
  hipLaunchKernelGGL(biharmonic_kernel<T>, dim3(tx * ty), dim3(512), 0, s, h, out, mh, ny, nx,
                     (T)(1.0 / dx), (T)(1.0 / dy), (T)kappa, tx, tx * ty);
  FVX_LAUNCH_CHECK();
  return kOk;
}

// ---------------------------------------------------------------------------
// Arakawa (1966) Jacobian (jacobian.py:22-131): interior-only output
// (batch, ny-2, nx-2); inputs (batch, ny, nx).  The three forms are summed in
// the reference's order and divided by 12 dx dy.
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T arakawa_point(const T* f, const T* g, int64_t c, int64_t ld) {
  const T fE = f[c + 1], fW = f[c - 1], fN = f[c + ld], fS = f[c - ld];
  const T fNE = f[c + ld + 1], fNW = f[c + ld - 1], fSE = f[c - ld + 1], fSW = f[c - ld - 1];
  const T gE = g[c + 1], gW = g[c - 1], gN = g[c + ld], gS = g[c - ld];
  const T gNE = g[c + ld + 1], gNW = g[c + ld - 1], gSE = g[c - ld + 1], gSW = g[c - ld - 1];
  const T jpp = (fE - fW) * (gN - gS) - (fN - fS) * (gE - gW);
  const T jpx = fE * (gNE - gSE) - fW * (gNW - gSW) - fN * (gNE - gNW) + fS * (gSE - gSW);
  const T jxp = gN * (fNE - fNW) - gS * (fSE - fSW) - gE * (fNE - fSE) + gW * (fNW - fSW);
  return (jpp + jpx) + jxp;
}

template <typename T>
__global__ __launch_bounds__(256) void arakawa_kernel(const T* __restrict__ f,
                                                      const T* __restrict__ g,
                                                      T* __restrict__ out, int ny, int nx,
                                                      T scale) {
  const int i = blockIdx.x * 64 + threadIdx.x;  // output column (0..nx-3)
  const int j = blockIdx.y * 4 + threadIdx.y;   // output row    (0..ny-3)
  const int b = blockIdx.z;
  if (i >= nx - 2 || j >= ny - 2) return;
  const int64_t base = (int64_t)b * ny * nx;
  const int64_t c = base + (int64_t)(j + 1) * nx + (i + 1);
  out[(int64_t)b * (ny - 2) * (nx - 2) + (int64_t)j * (nx - 2) + i] =
      arakawa_point(f, g, c, (int64_t)nx) / scale;
}

template <typename T>
int arakawa(const T* f, const T* g, T* out, int batch, int ny, int nx, double dx, double dy,
            hipStream_t s) {
  FVX_CHECK_ARG(f && g && out, "null field pointer");
  FVX_CHECK_ARG(batch >= 1 && batch <= 65535, "batch must be in [1, 65535], got %d", batch);
  FVX_CHECK_ARG(ny >= 3 && nx >= 3, "array too small (%d x %d)", ny, nx);
  dim3 blk(64, 4), grd(cdiv(nx - 2, 64), cdiv(ny - 2, 4), batch);
  hipLaunchKernelGGL(arakawa_kernel<T>, grd, blk, 0, s, f, g, out, ny, nx, (T)(12.0 * dx * dy));
  FVX_LAUNCH_CHECK();
  return kOk;
}

// ---------------------------------------------------------------------------
// masked_laplacian (iterative.py:144-191): Dirichlet outside the mask with a
// periodic roll at the array edge; the mask is a field of the same dtype.
// ---------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void masked_laplacian_kernel(
    const T* __restrict__ psi, const T* __restrict__ mask, T* __restrict__ out, int ny,
    int nx, T idx2, T idy2, T lam) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int ie = (i + 1 == nx) ? 0 : i + 1, iw = (i == 0) ? nx - 1 : i - 1;
  const int jn = (j + 1 == ny) ? 0 : j + 1, js = (j == 0) ? ny - 1 : j - 1;
  auto pm = [&](int jj, int ii) {
    const int64_t k = (int64_t)jj * nx + ii;
    return psi[k] * mask[k];
  };
  const T p = pm(j, i);
  const T lap = (pm(j, iw) + pm(j, ie) - T(2) * p) * idx2 + (pm(js, i) + pm(jn, i) - T(2) * p) * idy2;
  out[(int64_t)j * nx + i] = (lap - lam * p) * mask[(int64_t)j * nx + i];
}

template <typename T>
int masked_laplacian(const T* psi, const T* mask, T* out, int ny, int nx, double dx,
                     double dy, double lam, hipStream_t s) {
  FVX_CHECK_ARG(psi && mask && out, "null field pointer");
  FVX_CHECK_ARG(ny >= 1 && nx >= 1, "empty array");
  dim3 blk(64, 4), grd(cdiv(nx, 64), cdiv(ny, 4));
  hipLaunchKernelGGL(masked_laplacian_kernel<T>, grd, blk, 0, s, psi, mask, out, ny, nx,
                     (T)(1.0 / (dx * dx)), (T)(1.0 / (dy * dy)), (T)lam);
  FVX_LAUNCH_CHECK();
  return kOk;
}

// ---------------------------------------------------------------------------
// Periodic ghost fill of width ng (boundary.py:52-78), in place, in one
// launch: every ghost cell copies the interior cell its row and column wrap
// to (the reference fills rows, then columns, so a corner takes the
// diagonally opposite interior value: the same cell).  Reads touch interior
// cells only and writes ghost cells only, so the threads do not race.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int periodic_wrap(int x, int n, int ng) {
  return x < ng ? x + n - 2 * ng : (x >= n - ng ? x - n + 2 * ng : x);
}

template <typename T>
__global__ __launch_bounds__(256) void periodic_fill_kernel(T* f, int ny, int nx, int ng) {
  // cells 0 .. 2 ng nx - 1: the ghost rows (full width); then the ghost
  // columns of the interior rows, 2 ng per row
  const int64_t nrow = 2LL * ng * nx, ncol = 2LL * ng * (ny - 2 * ng);
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int j, i;
  if (k < nrow) {
    const int g = (int)(k / nx);
    i = (int)(k - (int64_t)g * nx);
    j = g < ng ? g : ny - 2 * ng + g;
  } else if (k < nrow + ncol) {
    const int64_t q = k - nrow;
    const int r = (int)(q / (2 * ng)), g = (int)(q - (int64_t)r * 2 * ng);
    j = ng + r;
    i = g < ng ? g : nx - 2 * ng + g;
  } else {
    return;
  }
  f[(int64_t)j * nx + i] = f[(int64_t)periodic_wrap(j, ny, ng) * nx + periodic_wrap(i, nx, ng)];
}

template <typename T>
int enforce_periodic(T* f, int ny, int nx, int ng, hipStream_t s) {
  FVX_CHECK_ARG(f, "null field pointer");
  FVX_CHECK_ARG(ng >= 1 && ny >= 3 * ng && nx >= 3 * ng,
                "ghost width %d too large for %d x %d", ng, ny, nx);
  const int64_t cells = 2LL * ng * nx + 2LL * ng * (ny - 2 * ng);
  hipLaunchKernelGGL(periodic_fill_kernel<T>, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0,
                     s, f, ny, nx, ng);
  FVX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace fvx

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

#define FVX_DT(SUF, T)                                                                      \
  int fvx_stencil2d_##SUF(int op, const T* a, const T* b, T* out, int ny, int nx,           \
                          double dx, double dy, const uint8_t* mask, void* stream) {        \
    return fvx::point_stencil<T>(op, a, b, out, ny, nx, dx, dy, mask, (hipStream_t)stream); \
  }                                                                                         \
  int fvx_diffusion2d_##SUF(const T* h, const T* kappa, double kappa0, T* out, int ny,      \
                            int nx, double dx, double dy, const uint8_t* mh,                \
                            const uint8_t* mu, const uint8_t* mv, void* stream) {           \
    return fvx::diffusion<T>(h, kappa, kappa0, out, ny, nx, dx, dy, mh, mu, mv,             \
                             (hipStream_t)stream);                                          \
  }                                                                                         \
  int fvx_diffusion2d_fluxes_##SUF(const T* h, const T* kappa, double kappa0, T* fx, T* fy, \
                                   int ny, int nx, double dx, double dy, const uint8_t* mu, \
                                   const uint8_t* mv, void* stream) {                       \
    return fvx::diffusion_fluxes<T>(h, kappa, kappa0, fx, fy, ny, nx, dx, dy, mu, mv,       \
                                    (hipStream_t)stream);                                   \
  }                                                                                         \
  int fvx_biharmonic2d_##SUF(const T* h, T* out, int ny, int nx, double dx, double dy,      \
                             double kappa, const uint8_t* mh, void* stream) {               \
    return fvx::biharmonic<T>(h, out, ny, nx, dx, dy, kappa, mh, (hipStream_t)stream);      \
  }                                                                                         \
  int fvx_arakawa_jacobian_##SUF(const T* f, const T* g, T* out, int batch, int ny, int nx, \
                                 double dx, double dy, void* stream) {                      \
    return fvx::arakawa<T>(f, g, out, batch, ny, nx, dx, dy, (hipStream_t)stream);          \
  }                                                                                         \
  int fvx_masked_laplacian_##SUF(const T* psi, const T* mask, T* out, int ny, int nx,       \
                                 double dx, double dy, double lambda_, void* stream) {      \
    return fvx::masked_laplacian<T>(psi, mask, out, ny, nx, dx, dy, lambda_,                \
                                    (hipStream_t)stream);                                   \
  }                                                                                         \
  int fvx_enforce_periodic_##SUF(T* field, int ny, int nx, int ng, void* stream) {          \
    return fvx::enforce_periodic<T>(field, ny, nx, ng, (hipStream_t)stream);                \
  }

FVX_DT(f32, float)
FVX_DT(f64, double)
#undef FVX_DT
}
