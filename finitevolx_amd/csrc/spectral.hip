// Spectral Helmholtz solvers (lap - lambda) psi = f on rectangles, the
// drop-in for solve_helmholtz_{dst,dct,fft} (re-exported from spectraldiffx
// v0.0.10 by solvers/spectral.py:14-128; used by scripts/qg_1p5_layer.py:443).
//
// Each transform diagonalises one of this library's 5-point operators
// exactly, so the solve is the operator's exact inverse:
//   dst : Dirichlet, zero just outside the array (DST-I; eigenvalues
//         -4/dx^2 sin^2(pi (k+1) / (2 (n+1))), dst1_eigenvalues)
//   dct : no-flux cell-centred Neumann (DCT-II / DCT-III pair; eigenvalues
//         -4/dx^2 sin^2(pi k / (2 n)); the k = l = 0 mode is set to zero
//         when the operator is singular)
//   fft : periodic (eigenvalues -4/dx^2 sin^2(pi k / n))
// Transforms along x run as batched real FFTs (hipFFT) of the odd / even
// extension of each row, written transposed so the y transform is again a
// batch of contiguous rows.

#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

#include "fvx_common.hpp"

namespace fvx {

enum SpecBC : int { kDst = 0, kDct = 1, kFft = 2 };

template <typename T>
struct Cplx;
template <>
struct Cplx<double> {
  using type = hipfftDoubleComplex;
};
template <>
struct Cplx<float> {
  using type = hipfftComplex;
};

// Row r of src (length n, pitch ld) -> extended real row of length M.
template <typename T>
__global__ void ext_build_kernel(const T* __restrict__ src, int64_t ld, T* __restrict__ ext,
                                 int rows, int n, int M, int bc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (k >= M || r >= rows) return;
  const T* x = src + (int64_t)r * ld;
  T v;
  if (bc == kDst) {  // [0, x0..x_{n-1}, 0, -x_{n-1}..-x0], M = 2(n+1)
    if (k == 0 || k == n + 1) v = T(0);
    else if (k <= n) v = x[k - 1];
    else v = -x[2 * n + 1 - k];
  } else {  // [x0..x_{n-1}, x_{n-1}..x0], M = 2n
    v = (k < n) ? x[k] : x[2 * n - 1 - k];
  }
  ext[(int64_t)r * M + k] = v;
}

// Spectrum row r (length M/2+1) -> transform coefficient k, written
// transposed: out[k * rows + r].
template <typename T, typename C>
__global__ void ext_extract_kernel(const C* __restrict__ spec, T* __restrict__ out, int rows,
                                   int n, int M, int bc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (k >= n || r >= rows) return;
  const C* row = spec + (int64_t)r * (M / 2 + 1);
  T v;
  if (bc == kDst) {
    v = T(-0.5) * (T)row[k + 1].y;  // sum_j x_j sin(pi (k+1) (j+1) / (n+1))
  } else {
    const double th = -M_PI * k / (2.0 * n);
    const double c = cos(th), s = sin(th);
    v = (T)(0.5 * ((double)row[k].x * c - (double)row[k].y * s));  // DCT-II
  }
  out[(int64_t)k * rows + r] = v;
}

// DCT-III via C2R: V_k = X_k e^{i pi k / 2n} (k < n), V_n = 0.
template <typename T, typename C>
__global__ void dct3_build_kernel(const T* __restrict__ X, C* __restrict__ V, int rows, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (k > n || r >= rows) return;
  C v;
  if (k == n) {
    v.x = 0;
    v.y = 0;
  } else {
    const double th = M_PI * k / (2.0 * n);
    const double x = (double)X[(int64_t)r * n + k];
    v.x = (T)(x * cos(th));
    v.y = (T)(x * sin(th));
  }
  V[(int64_t)r * (n + 1) + k] = v;
}

// First n samples of each length-2n C2R row, scaled, written transposed.
template <typename T>
__global__ void dct3_extract_kernel(const T* __restrict__ v, T* __restrict__ out, int rows, int n,
                                    T scale) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (k >= n || r >= rows) return;
  out[(int64_t)k * rows + r] = v[(int64_t)r * 2 * n + k] * scale;
}

// Divide the (nx, ny)-transposed coefficients by the operator eigenvalues.
template <typename T>
__global__ void eig_divide_kernel(T* __restrict__ c, int nx, int ny, int bc, double dx, double dy,
                                  double lam, double scale) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;  // y index (fast)
  const int k = blockIdx.y;                             // x index
  if (l >= ny || k >= nx) return;
  double ex, ey;
  if (bc == kDst) {
    const double sx = sin(M_PI * (k + 1) / (2.0 * (nx + 1))), sy = sin(M_PI * (l + 1) / (2.0 * (ny + 1)));
    ex = -4.0 / (dx * dx) * sx * sx;
    ey = -4.0 / (dy * dy) * sy * sy;
  } else {
    const double sx = sin(M_PI * k / (2.0 * nx)), sy = sin(M_PI * l / (2.0 * ny));
    ex = -4.0 / (dx * dx) * sx * sx;
    ey = -4.0 / (dy * dy) * sy * sy;
  }
  const double d = ex + ey - lam;
  const int64_t i = (int64_t)k * ny + l;
  c[i] = (d != 0.0) ? (T)((double)c[i] * scale / d) : T(0);
}

template <typename C>
__global__ void fft_divide_kernel(C* __restrict__ c, int ny, int nx, double dx, double dy,
                                  double lam, double scale) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;  // x wavenumber, 0..nx/2
  const int l = blockIdx.y;                             // y wavenumber, 0..ny-1
  const int nk = nx / 2 + 1;
  if (k >= nk || l >= ny) return;
  const double sx = sin(M_PI * k / nx), sy = sin(M_PI * l / ny);
  const double d = -4.0 / (dx * dx) * sx * sx - 4.0 / (dy * dy) * sy * sy - lam;
  C& v = c[(int64_t)l * nk + k];
  const double f = (d != 0.0) ? scale / d : 0.0;
  v.x = v.x * f;
  v.y = v.y * f;
}

template <typename T>
struct FftTypes;
template <>
struct FftTypes<double> {
  static constexpr hipfftType r2c = HIPFFT_D2Z, c2r = HIPFFT_Z2D;
  static hipfftResult fwd(hipfftHandle p, double* i, hipfftDoubleComplex* o) {
    return hipfftExecD2Z(p, i, o);
  }
  static hipfftResult inv(hipfftHandle p, hipfftDoubleComplex* i, double* o) {
    return hipfftExecZ2D(p, i, o);
  }
};
template <>
struct FftTypes<float> {
  static constexpr hipfftType r2c = HIPFFT_R2C, c2r = HIPFFT_C2R;
  static hipfftResult fwd(hipfftHandle p, float* i, hipfftComplex* o) {
    return hipfftExecR2C(p, i, o);
  }
  static hipfftResult inv(hipfftHandle p, hipfftComplex* i, float* o) {
    return hipfftExecC2R(p, i, o);
  }
};

struct SpecPlan {
  int ny, nx, bc, prec;
  // per-direction batched 1-D plans (x: ny rows of length Mx; y: nx rows of length My)
  hipfftHandle fx = 0, fy = 0, ix = 0, iy = 0, f2 = 0, i2 = 0;
  void* ext = nullptr;   // real extension buffer
  void* spec = nullptr;  // complex buffer
  void* a = nullptr;     // transposed real (nx, ny)
  void* b = nullptr;     // real (ny, nx) / second transposed buffer
};

#define FVX_FFT(call)                                                  \
  do {                                                                 \
    hipfftResult r_ = (call);                                          \
    if (r_ != HIPFFT_SUCCESS) {                                        \
      ::fvx::set_error("hipFFT error %d at %s:%d", (int)r_, __FILE__, __LINE__); \
      return ::fvx::kHipError;                                         \
    }                                                                  \
  } while (0)

template <typename T>
static int spec_create(SpecPlan* P) {
  const int ny = P->ny, nx = P->nx;
  using FT = FftTypes<T>;
  using C = typename Cplx<T>::type;
  size_t ext_elems = 0, spec_elems = 0;
  if (P->bc == kFft) {
    FVX_FFT(hipfftPlan2d(&P->f2, ny, nx, FT::r2c));
    FVX_FFT(hipfftPlan2d(&P->i2, ny, nx, FT::c2r));
    spec_elems = (size_t)ny * (nx / 2 + 1);
  } else {
    const int Mx = P->bc == kDst ? 2 * (nx + 1) : 2 * nx;
    const int My = P->bc == kDst ? 2 * (ny + 1) : 2 * ny;
    FVX_FFT(hipfftPlan1d(&P->fx, Mx, FT::r2c, ny));
    FVX_FFT(hipfftPlan1d(&P->fy, My, FT::r2c, nx));
    if (P->bc == kDct) {
      FVX_FFT(hipfftPlan1d(&P->ix, 2 * nx, FT::c2r, ny));
      FVX_FFT(hipfftPlan1d(&P->iy, 2 * ny, FT::c2r, nx));
    }
    ext_elems = std::max((size_t)ny * Mx, (size_t)nx * My);
    spec_elems = std::max((size_t)ny * (Mx / 2 + 1), (size_t)nx * (My / 2 + 1));
  }
  const size_t N = (size_t)ny * nx;
  FVX_HIP(hipMalloc(&P->spec, spec_elems * sizeof(C)));
  if (ext_elems) FVX_HIP(hipMalloc(&P->ext, ext_elems * sizeof(T)));
  FVX_HIP(hipMalloc(&P->a, N * sizeof(T)));
  FVX_HIP(hipMalloc(&P->b, N * sizeof(T)));
  return kOk;
}

// Forward sine/cosine transform of `rows` rows of length n (pitch ld) ->
// transposed coefficients (n, rows) in out.
template <typename T>
static int fwd_rows(SpecPlan* P, hipfftHandle plan, const T* src, int64_t ld, int rows, int n,
                    T* out, hipStream_t s) {
  using C = typename Cplx<T>::type;
  const int M = P->bc == kDst ? 2 * (n + 1) : 2 * n;
  hipLaunchKernelGGL(ext_build_kernel<T>, dim3(cdiv(M, 256), rows), dim3(256), 0, s, src, ld,
                     (T*)P->ext, rows, n, M, P->bc);
  FVX_FFT(hipfftSetStream(plan, s));
  FVX_FFT(FftTypes<T>::fwd(plan, (T*)P->ext, (C*)P->spec));
  hipLaunchKernelGGL((ext_extract_kernel<T, C>), dim3(cdiv(n, 256), rows), dim3(256), 0, s,
                     (const C*)P->spec, out, rows, n, M, P->bc);
  FVX_LAUNCH_CHECK();
  return kOk;
}

// Inverse along rows (DST-I is its own inverse up to 2/(n+1); DCT-III via
// C2R): coefficients (rows, n) -> transposed values (n, rows), scaled.
template <typename T>
static int inv_rows(SpecPlan* P, hipfftHandle fplan, hipfftHandle iplan, const T* src, int rows,
                    int n, T* out, hipStream_t s) {
  using C = typename Cplx<T>::type;
  if (P->bc == kDst) {
    int rc = fwd_rows<T>(P, fplan, src, n, rows, n, out, s);
    return rc;  // scale folded into the eigenvalue division
  }
  hipLaunchKernelGGL((dct3_build_kernel<T, C>), dim3(cdiv(n + 1, 256), rows), dim3(256), 0, s,
                     src, (C*)P->spec, rows, n);
  FVX_FFT(hipfftSetStream(iplan, s));
  FVX_FFT(FftTypes<T>::inv(iplan, (C*)P->spec, (T*)P->ext));
  hipLaunchKernelGGL(dct3_extract_kernel<T>, dim3(cdiv(n, 256), rows), dim3(256), 0, s,
                     (const T*)P->ext, out, rows, n, (T)(1.0 / n));
  FVX_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
__global__ void copy_pitched_kernel(const T* __restrict__ src, int64_t lds, T* __restrict__ dst,
                                    int64_t ldd, int ny, int nx) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i < nx && j < ny) dst[(int64_t)j * ldd + i] = src[(int64_t)j * lds + i];
}

template <typename T>
static int spec_solve(SpecPlan* P, const T* rhs, int64_t ld_rhs, T* out, int64_t ld_out,
                      double dx, double dy, double lam, hipStream_t s) {
  using C = typename Cplx<T>::type;
  const int ny = P->ny, nx = P->nx;
  T* A = (T*)P->a;
  T* B = (T*)P->b;
  if (P->bc == kFft) {
    hipLaunchKernelGGL(copy_pitched_kernel<T>, dim3(cdiv(nx, 64), cdiv(ny, 4)), dim3(64, 4), 0, s,
                       rhs, ld_rhs, A, (int64_t)nx, ny, nx);
    FVX_FFT(hipfftSetStream(P->f2, s));
    FVX_FFT(FftTypes<T>::fwd(P->f2, A, (C*)P->spec));
    hipLaunchKernelGGL((fft_divide_kernel<C>), dim3(cdiv(nx / 2 + 1, 256), ny), dim3(256), 0, s,
                       (C*)P->spec, ny, nx, dx, dy, lam, 1.0 / ((double)nx * ny));
    FVX_FFT(hipfftSetStream(P->i2, s));
    FVX_FFT(FftTypes<T>::inv(P->i2, (C*)P->spec, B));
    hipLaunchKernelGGL(copy_pitched_kernel<T>, dim3(cdiv(nx, 64), cdiv(ny, 4)), dim3(64, 4), 0, s,
                       B, (int64_t)nx, out, ld_out, ny, nx);
    FVX_LAUNCH_CHECK();
    return kOk;
  }
  // x transform of the ny rows -> A (nx, ny); y transform of the nx rows -> B (ny, nx)
  int rc = fwd_rows<T>(P, P->fx, rhs, ld_rhs, ny, nx, A, s);
  if (!rc) rc = fwd_rows<T>(P, P->fy, A, ny, nx, ny, B, s);
  if (rc) return rc;
  // B is (ny, nx) in (l, k) order; divide with the eigenvalues (B viewed as
  // (ny rows of nx): index l * nx + k).  Reuse eig_divide with swapped roles.
  double scale = 1.0;
  if (P->bc == kDst) scale = (2.0 / (nx + 1)) * (2.0 / (ny + 1));
  hipLaunchKernelGGL(eig_divide_kernel<T>, dim3(cdiv(nx, 256), ny), dim3(256), 0, s, B, ny, nx,
                     P->bc, dy, dx, lam, scale);
  // inverse along y (rows of length ny of the (nx, ny) layout) and x.
  rc = inv_rows<T>(P, P->fx, P->ix, B, ny, nx, A, s);  // rows = ny (l), length nx (k) -> A (nx, ny)
  if (!rc) rc = inv_rows<T>(P, P->fy, P->iy, A, nx, ny, B, s);  // -> B (ny, nx)
  if (rc) return rc;
  hipLaunchKernelGGL(copy_pitched_kernel<T>, dim3(cdiv(nx, 64), cdiv(ny, 4)), dim3(64, 4), 0, s,
                     B, (int64_t)nx, out, ld_out, ny, nx);
  FVX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace fvx

extern "C" {

// bc: 0 = DST-I (Dirichlet), 1 = DCT-II (Neumann), 2 = FFT (periodic);
// precision 32 or 64.  ny, nx: the array the solve runs on (interior only).
int fvx_spectral_plan_create(int ny, int nx, int bc, int precision, void** plan_out) {
  using namespace fvx;
  FVX_CHECK_ARG(plan_out, "null plan_out");
  FVX_CHECK_ARG(ny >= 2 && nx >= 2, "spectral solve needs at least 2 x 2");
  FVX_CHECK_ARG(bc >= 0 && bc <= 2, "bc must be 0 (dst), 1 (dct) or 2 (fft)");
  FVX_CHECK_ARG(precision == 32 || precision == 64, "precision must be 32 or 64");
  SpecPlan* P = new SpecPlan();
  P->ny = ny;
  P->nx = nx;
  P->bc = bc;
  P->prec = precision;
  int rc = precision == 64 ? spec_create<double>(P) : spec_create<float>(P);
  if (rc) {
    delete P;
    return rc;
  }
  *plan_out = P;
  return kOk;
}

int fvx_spectral_plan_destroy(void* plan) {
  using namespace fvx;
  if (!plan) return kOk;
  SpecPlan* P = (SpecPlan*)plan;
  (void)hipDeviceSynchronize();
  for (hipfftHandle h : {P->fx, P->fy, P->ix, P->iy, P->f2, P->i2})
    if (h) hipfftDestroy(h);
  (void)hipFree(P->ext);
  (void)hipFree(P->spec);
  (void)hipFree(P->a);
  (void)hipFree(P->b);
  delete P;
  return kOk;
}

// (lap - lambda) out = rhs for the plan's boundary condition; rhs / out pitched.
int fvx_spectral_helmholtz_f64(void* plan, const double* rhs, int64_t ld_rhs, double* out,
                               int64_t ld_out, double dx, double dy, double lambda_,
                               void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(plan && rhs && out, "null argument");
  SpecPlan* P = (SpecPlan*)plan;
  FVX_CHECK_ARG(P->prec == 64, "plan was built for float32");
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  return spec_solve<double>(P, rhs, ld_rhs, out, ld_out, dx, dy, lambda_, (hipStream_t)stream);
}
int fvx_spectral_helmholtz_f32(void* plan, const float* rhs, int64_t ld_rhs, float* out,
                               int64_t ld_out, double dx, double dy, double lambda_,
                               void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(plan && rhs && out, "null argument");
  SpecPlan* P = (SpecPlan*)plan;
  FVX_CHECK_ARG(P->prec == 32, "plan was built for float64");
  FVX_CHECK_ARG(dx > 0 && dy > 0, "grid spacing must be positive");
  return spec_solve<float>(P, rhs, ld_rhs, out, ld_out, dx, dy, lambda_, (hipStream_t)stream);
}
}
