// Elementwise vector kernels: the Runge-Kutta stage combination used by
// rk3_ssp_step (timestepping/explicit_rk.py:104-121) and helpers.

#include "fvx_common.hpp"

namespace fvx {

// out = a*x + b*y + c*z   (y or z may be null -> term dropped)
template <typename T>
__global__ __launch_bounds__(256) void lincomb3_kernel(int64_t n, T a, const T* __restrict__ x,
                                                       T b, const T* __restrict__ y, T c,
                                                       const T* __restrict__ z,
                                                       T* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += stride) {
    T r = a * x[k];
    if (y) r += b * y[k];
    if (z) r += c * z[k];
    out[k] = r;
  }
}

template <typename T>
int lincomb3(int64_t n, double a, const T* x, double b, const T* y, double c, const T* z,
             T* out, hipStream_t s) {
  FVX_CHECK_ARG(x && out, "null pointer");
  FVX_CHECK_ARG(n >= 0, "negative length");
  if (n == 0) return kOk;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(lincomb3_kernel<T>, dim3(blocks), dim3(256), 0, s, n, (T)a, x, (T)b, y,
                     (T)c, z, out);
  FVX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace fvx

extern "C" {
int fvx_lincomb3_f64(int64_t n, double a, const double* x, double b, const double* y, double c,
                     const double* z, double* out, void* stream) {
  return fvx::lincomb3<double>(n, a, x, b, y, c, z, out, (hipStream_t)stream);
}
int fvx_lincomb3_f32(int64_t n, double a, const float* x, double b, const float* y, double c,
                     const float* z, float* out, void* stream) {
  return fvx::lincomb3<float>(n, a, x, b, y, c, z, out, (hipStream_t)stream);
}
}
