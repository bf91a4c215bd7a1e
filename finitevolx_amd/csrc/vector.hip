// Elementwise vector kernels: the Runge-Kutta stage combination used by
// rk3_ssp_step (timestepping/explicit_rk.py:104-121) and helpers.

#include "fvx_common.hpp"

namespace fvx {

// out = a*x + b*y + c*z + d*w   (y, z or w may be null -> term dropped).
// out may be one of the inputs: each element is read before it is written.
template <typename T>
__global__ __launch_bounds__(256) void lincomb4_kernel(int64_t n, T a, const T* x, T b,
                                                       const T* y, T c, const T* z, T d,
                                                       const T* w, T* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += stride) {
    T r = a * x[k];
    if (y) r += b * y[k];
    if (z) r += c * z[k];
    if (w) r += d * w[k];
    out[k] = r;
  }
}

// The same over 16-byte vectors (two doubles / four floats per lane) when
// every array is 16-byte aligned; the last n % W elements by the first lanes.
template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int W = 2;
};
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int W = 4;
};

template <typename T>
__global__ __launch_bounds__(256) void lincomb4_vec_kernel(int64_t n, T a, const T* x, T b,
                                                           const T* y, T c, const T* z, T d,
                                                           const T* w, T* out) {
  using V = typename Vec16<T>::type;
  constexpr int W = Vec16<T>::W;
  const int64_t nv = n / W;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t k = t0; k < nv; k += stride) {
    V r = a * ((const V*)x)[k];
    if (y) r += b * ((const V*)y)[k];
    if (z) r += c * ((const V*)z)[k];
    if (w) r += d * ((const V*)w)[k];
    ((V*)out)[k] = r;
  }
  for (int64_t k = nv * W + t0; k < n; k += stride) {
    T r = a * x[k];
    if (y) r += b * y[k];
    if (z) r += c * z[k];
    if (w) r += d * w[k];
    out[k] = r;
  }
}

template <typename T>
int lincomb4(int64_t n, double a, const T* x, double b, const T* y, double c, const T* z,
             double d, const T* w, T* out, hipStream_t s) {
  FVX_CHECK_ARG(x && out, "null pointer");
  FVX_CHECK_ARG(n >= 0, "negative length");
  if (n == 0) return kOk;
  auto al = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  if (al(x) && al(y) && al(z) && al(w) && al(out)) {
    const int64_t nv = n / Vec16<T>::W;
    const int blocks = (int)std::min<int64_t>(std::max<int64_t>((nv + 255) / 256, 1), 8192);
    hipLaunchKernelGGL(lincomb4_vec_kernel<T>, dim3(blocks), dim3(256), 0, s, n, (T)a, x, (T)b,
                       y, (T)c, z, (T)d, w, out);
  } else {
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(lincomb4_kernel<T>, dim3(blocks), dim3(256), 0, s, n, (T)a, x, (T)b, y,
                       (T)c, z, (T)d, w, out);
  }
  FVX_LAUNCH_CHECK();
  return kOk;
}
template <typename T>
int lincomb3(int64_t n, double a, const T* x, double b, const T* y, double c, const T* z,
             T* out, hipStream_t s) {
  return lincomb4<T>(n, a, x, b, y, c, z, 0.0, nullptr, out, s);
}
template int lincomb3<float>(int64_t, double, const float*, double, const float*, double,
                             const float*, float*, hipStream_t);
template int lincomb3<double>(int64_t, double, const double*, double, const double*, double,
                              const double*, double*, hipStream_t);
template int lincomb4<float>(int64_t, double, const float*, double, const float*, double,
                             const float*, double, const float*, float*, hipStream_t);
template int lincomb4<double>(int64_t, double, const double*, double, const double*, double,
                              const double*, double, const double*, double*, hipStream_t);

// out[k, p] = sum_l C[k, l] in[l, p] for nl <= 8 layers (layer_to_mode /
// mode_to_layer, vertical/vertical_modes.py:193-263): one pass, each
// thread owns one horizontal point and all its layers.
struct Mix8 {
  double c[64];
};

template <typename T>
__global__ __launch_bounds__(256) void layer_mix_kernel(const T* __restrict__ in,
                                                        T* __restrict__ out, Mix8 C, int nl,
                                                        int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += stride) {
    T v[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) v[l] = (l < nl) ? in[(int64_t)l * n + p] : T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k >= nl) break;
      double acc = 0.0;
#pragma unroll
      for (int l = 0; l < 8; ++l)
        if (l < nl) acc += C.c[k * 8 + l] * (double)v[l];
      out[(int64_t)k * n + p] = (T)acc;
    }
  }
}

template <typename T>
int layer_mix(const T* in, T* out, const double* C, int nl, int64_t n, hipStream_t s) {
  FVX_CHECK_ARG(in && out && C, "null pointer");
  FVX_CHECK_ARG(nl >= 1 && nl <= 8, "1 <= nl <= 8 layers supported, got %d", nl);
  FVX_CHECK_ARG(in != out, "in-place layer transform is not supported");
  Mix8 m;
  for (int k = 0; k < 64; ++k) m.c[k] = 0.0;
  for (int k = 0; k < nl; ++k)
    for (int l = 0; l < nl; ++l) m.c[k * 8 + l] = C[k * nl + l];
  if (n == 0) return kOk;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(layer_mix_kernel<T>, dim3(blocks), dim3(256), 0, s, in, out, m, nl, n);
  FVX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace fvx

extern "C" {
// C: host row-major nl x nl matrix; in / out: device (nl, n).
int fvx_layer_mix_f64(const double* in, double* out, const double* C, int nl, int64_t n,
                      void* stream) {
  return fvx::layer_mix<double>(in, out, C, nl, n, (hipStream_t)stream);
}
int fvx_layer_mix_f32(const float* in, float* out, const double* C, int nl, int64_t n,
                      void* stream) {
  return fvx::layer_mix<float>(in, out, C, nl, n, (hipStream_t)stream);
}
int fvx_lincomb3_f64(int64_t n, double a, const double* x, double b, const double* y, double c,
                     const double* z, double* out, void* stream) {
  return fvx::lincomb3<double>(n, a, x, b, y, c, z, out, (hipStream_t)stream);
}
int fvx_lincomb3_f32(int64_t n, double a, const float* x, double b, const float* y, double c,
                     const float* z, float* out, void* stream) {
  return fvx::lincomb3<float>(n, a, x, b, y, c, z, out, (hipStream_t)stream);
}
int fvx_lincomb4_f64(int64_t n, double a, const double* x, double b, const double* y, double c,
                     const double* z, double d, const double* w, double* out, void* stream) {
  return fvx::lincomb4<double>(n, a, x, b, y, c, z, d, w, out, (hipStream_t)stream);
}
int fvx_lincomb4_f32(int64_t n, double a, const float* x, double b, const float* y, double c,
                     const float* z, double d, const float* w, float* out, void* stream) {
  return fvx::lincomb4<float>(n, a, x, b, y, c, z, d, w, out, (hipStream_t)stream);
}
}
