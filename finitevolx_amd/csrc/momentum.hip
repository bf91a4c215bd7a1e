// Vector-invariant momentum advection (MomentumAdvection2D.__call__,
// diffusion/momentum.py:37-248): du = +(zeta v)_u - dK/dx at U points,
// dv = -(zeta u)_v - dK/dy at V points, written on [2:-2, 2:-2] (zero
// elsewhere), with zeta = curl(u, v) at X points, K = (u_T^2 + v_T^2) / 2 at
// T points and the vorticity flux in Sadourny's energy or enstrophy form or
// the Arakawa-Lamb blend (1/3 energy + 2/3 enstrophy).  With a mask every
// intermediate is multiplied by the mask of its stagger, as the reference's
// masked Difference2D / Interpolation2D do, and the result by mask.u / mask.v.
//
// One thread per output point computes the whole chain from its u, v
// neighbourhood (rows j-1..j+1, columns i-1..i+1): no intermediate array
// reaches HBM; traffic per cell is read u, v, write du, dv (+ 4 mask bytes).

#include "fvx_common.hpp"

namespace fvx {

enum MomentumScheme : int { kEnergy = 0, kEnstrophy = 1, kArakawaLamb = 2 };

template <typename T, int SCHEME, bool MASKED>
__global__ __launch_bounds__(256) void momentum_kernel(
    const T* __restrict__ u, const T* __restrict__ v, T* __restrict__ du, T* __restrict__ dv,
    const uint8_t* __restrict__ mh, const uint8_t* __restrict__ mu,
    const uint8_t* __restrict__ mv, const uint8_t* __restrict__ mq, int ny, int nx, T idx,
    T idy) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= nx || j >= ny) return;
  const int64_t c = (int64_t)j * nx + i;
  if (j < 2 || j > ny - 3 || i < 2 || i > nx - 3) {
    du[c] = T(0);
    dv[c] = T(0);
    return;
  }
  auto at = [&](const T* a, int dj, int di) { return a[c + (int64_t)dj * nx + di]; };
  auto m = [&](const uint8_t* a, int dj, int di) -> T {
    if constexpr (MASKED) return a[c + (int64_t)dj * nx + di] ? T(1) : T(0);
    return T(1);
  };
  // curl at the X point (j+dj, i+di): interior rows/columns only (all the
  // points read below are interior for 2 <= j <= ny-3, 2 <= i <= nx-3).
  auto zeta = [&](int dj, int di) {
    return ((at(v, dj, di + 1) - at(v, dj, di)) * idx - (at(u, dj + 1, di) - at(u, dj, di)) * idy) *
           m(mq, dj, di);
  };
  // kinetic energy at the T point (j+dj, i+di)
  auto K = [&](int dj, int di) {
    const T ut = T(0.5) * (at(u, dj, di) + at(u, dj, di - 1)) * m(mh, dj, di);
    const T vt = T(0.5) * (at(v, dj, di) + at(v, dj - 1, di)) * m(mh, dj, di);
    return T(0.5) * (ut * ut + vt * vt);
  };
  const T mu0 = m(mu, 0, 0), mv0 = m(mv, 0, 0);
  const T z00 = zeta(0, 0), zS = zeta(-1, 0), zW = zeta(0, -1);
  const T K00 = K(0, 0);
  const T dKdx = (K(0, 1) - K00) * idx * mu0;
  const T dKdy = (K(1, 0) - K00) * idy * mv0;
  T zvu = T(0), zuv = T(0);
  if constexpr (SCHEME == kEnergy || SCHEME == kArakawaLamb) {
    const T z_u = T(0.5) * (z00 + zS) * mu0;
    const T v_u = T(0.25) * (at(v, 0, 0) + at(v, -1, 0) + at(v, 0, 1) + at(v, -1, 1)) * mu0;
    const T z_v = T(0.5) * (z00 + zW) * mv0;
    const T u_v = T(0.25) * (at(u, 0, 0) + at(u, 1, 0) + at(u, 0, -1) + at(u, 1, -1)) * mv0;
    zvu = z_u * v_u;
    zuv = z_v * u_v;
  }
  if constexpr (SCHEME == kEnstrophy || SCHEME == kArakawaLamb) {
    auto vq = [&](int dj, int di) {
      return T(0.5) * (at(v, dj, di) + at(v, dj, di + 1)) * m(mq, dj, di);
    };
    auto uq = [&](int dj, int di) {
      return T(0.5) * (at(u, dj, di) + at(u, dj + 1, di)) * m(mq, dj, di);
    };
    const T zs_u = T(0.5) * (z00 * vq(0, 0) + zS * vq(-1, 0)) * mu0;
    const T zs_v = T(0.5) * (z00 * uq(0, 0) + zW * uq(0, -1)) * mv0;
    if constexpr (SCHEME == kEnstrophy) {
      zvu = zs_u;
      zuv = zs_v;
    } else {
      constexpr T a = T(1.0 / 3.0);
      zvu = a * zvu + (T(1) - a) * zs_u;
      zuv = a * zuv + (T(1) - a) * zs_v;
    }
  }
  du[c] = (zvu - dKdx) * mu0;
  dv[c] = (-zuv - dKdy) * mv0;
}

template <typename T>
int momentum(const T* u, const T* v, T* du, T* dv, int scheme, const uint8_t* mh,
             const uint8_t* mu, const uint8_t* mv, const uint8_t* mq, int ny, int nx, double dx,
             double dy, hipStream_t s) {
  FVX_CHECK_ARG(u && v && du && dv, "null field pointer");
  FVX_CHECK_ARG(ny >= 3 && nx >= 3, "array too small (%d x %d)", ny, nx);
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  FVX_CHECK_ARG(scheme >= 0 && scheme <= 2,
                "Unknown scheme %d: 0 energy, 1 enstrophy, 2 Arakawa-Lamb", scheme);
  const bool masked = mh || mu || mv || mq;
  FVX_CHECK_ARG(!masked || (mh && mu && mv && mq), "pass all four stagger masks or none");
  const dim3 g(cdiv(nx, 64), cdiv(ny, 4)), b(256);
  const T ix = (T)(1.0 / dx), iy = (T)(1.0 / dy);
#define FVX_MOM(S, M)                                                                             \
  hipLaunchKernelGGL((momentum_kernel<T, S, M>), g, b, 0, s, u, v, du, dv, mh, mu, mv, mq, ny, nx, \
                     ix, iy)
  if (masked) {
    if (scheme == kEnergy) FVX_MOM(kEnergy, true);
    else if (scheme == kEnstrophy) FVX_MOM(kEnstrophy, true);
    else FVX_MOM(kArakawaLamb, true);
  } else {
    if (scheme == kEnergy) FVX_MOM(kEnergy, false);
    else if (scheme == kEnstrophy) FVX_MOM(kEnstrophy, false);
    else FVX_MOM(kArakawaLamb, false);
  }
#undef FVX_MOM
  FVX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace fvx

extern "C" {
int fvx_momentum_advection2d_f32(const float* u, const float* v, float* du, float* dv, int scheme,
                                 const uint8_t* mask_h, const uint8_t* mask_u,
                                 const uint8_t* mask_v, const uint8_t* mask_xy, int ny, int nx,
                                 double dx, double dy, void* stream) {
  return fvx::momentum<float>(u, v, du, dv, scheme, mask_h, mask_u, mask_v, mask_xy, ny, nx, dx,
                              dy, (hipStream_t)stream);
}
int fvx_momentum_advection2d_f64(const double* u, const double* v, double* du, double* dv,
                                 int scheme, const uint8_t* mask_h, const uint8_t* mask_u,
                                 const uint8_t* mask_v, const uint8_t* mask_xy, int ny, int nx,
                                 double dx, double dy, void* stream) {
  return fvx::momentum<double>(u, v, du, dv, scheme, mask_h, mask_u, mask_v, mask_xy, ny, nx, dx,
                               dy, (hipStream_t)stream);
}
}
