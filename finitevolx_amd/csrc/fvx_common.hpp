// Shared helpers for the finitevolx HIP library (gfx950 / CDNA4).
//
// Conventions used by every kernel in this library:
//  * Fields are row-major [ny][ld] arrays of T (float or double); ``ld`` is the
//    row pitch in elements (>= nx).  Operators that mirror the reference's
//    ghost-ring contract take the FULL array (ghost ring included) and write a
//    zero ghost ring.
//  * Masks are uint8 arrays (0 = dry, nonzero = wet) with the same pitch as
//    the field they gate, or nullptr for "all wet".
//  * Every entry point takes a hipStream_t and only enqueues work; it returns
//    0 on success and a negative status otherwise (message in fvx_last_error).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace fvx {

enum Status : int {
  kOk = 0,
  kInvalidArgument = -1,
  kHipError = -2,
  kNotConverged = 1,  // informational; never returned by kernels
};

// Thread-local last error message (read through fvx_last_error()).
void set_error(const char* fmt, ...);
const char* last_error();

#define FVX_CHECK_ARG(cond, ...)              \
  do {                                        \
    if (!(cond)) {                            \
      ::fvx::set_error(__VA_ARGS__);          \
      return ::fvx::kInvalidArgument;         \
    }                                         \
  } while (0)

#define FVX_HIP(call)                                                          \
  do {                                                                         \
    hipError_t err_ = (call);                                                  \
    if (err_ != hipSuccess) {                                                  \
      ::fvx::set_error("HIP error %s at %s:%d", hipGetErrorString(err_),       \
                       __FILE__, __LINE__);                                    \
      return ::fvx::kHipError;                                                 \
    }                                                                          \
  } while (0)

#define FVX_LAUNCH_CHECK() FVX_HIP(hipGetLastError())

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// Map a linear workgroup id so that consecutive tiles land on the same XCD.
// The hardware dispatches workgroup b to XCD (b % 8); tiles that share halo
// rows should share an L2, so tile t is placed on XCD t / (n / 8).
__device__ __forceinline__ int xcd_swizzle(int b, int n) {
  constexpr int kXcd = 8;
  if (n < 2 * kXcd) return b;
  const int per = n / kXcd;       // full tiles per XCD
  const int full = per * kXcd;
  if (b >= full) return b;         // ragged tail keeps its id
  return (b % kXcd) * per + (b / kXcd);
}

__device__ __forceinline__ float wet(const uint8_t* m, int64_t idx) {
  return m ? (m[idx] ? 1.0f : 0.0f) : 1.0f;
}

// Value of lane + 1 (wave_shl 1) and lane - 1 (wave_shr 1); 0 where there is
// none (lane 63, lane 0).
__device__ __forceinline__ float lane_east(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_west(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

// Buffer accesses: a 4-register descriptor per array (base, byte size), a
// wave-uniform row offset in a scalar register and the lane's column offset,
// so no per-access 64-bit address arithmetic; accesses past the array's end
// read 0 / are dropped.  Arrays must be under 2^31 bytes.  The last
// descriptor word, 0x00020000, is the CDNA (gfx9-family) setting for raw
// 32-bit data (data format 32, no swizzle, no index stride).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const float* p, int n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, n * 4, 0x00020000);
}
__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, unsigned col, int row_off) {
  return __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, col * 4u, row_off * 4, 0));
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, unsigned col, int row_off,
                                          float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_int(v), r, col * 4u, row_off * 4, 0);
}

// DPP lane moves; lanes without a source read 0 (bound_ctrl), so the sums
// below need no initialised destination.
template <int kCtrl>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, kCtrl, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), kCtrl, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int kCtrl>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kCtrl, 0xf, 0xf, true));
}

// Sum over the 64 lanes of a wavefront with DPP lane moves (VALU only, no
// LDS round trips); the total is valid in lane 63.  Pairs and quads
// (quad_perm), rows of 16 (row_shr 4, 8: lane 15 of each row holds the row
// sum), then rows 0+1 into lane 31 and rows 2+3 into lane 63 (row_bcast 15)
// and lane 31 into lane 63 (row_bcast 31).  Other lanes hold partial sums.
template <typename T>
__device__ __forceinline__ T wave_sum_last(T v) {
  v += dpp_move<0xb1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_move<0x4e>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_move<0x114>(v);  // row_shr 4
  v += dpp_move<0x118>(v);  // row_shr 8
  v += dpp_move<0x142>(v);  // row_bcast 15
  v += dpp_move<0x143>(v);  // row_bcast 31
  return v;
}

// Wavefront (64-lane) and workgroup reductions in double.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <int kThreads>
__device__ __forceinline__ double block_sum(double v, double* smem) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) s += smem[w];
  }
  __syncthreads();
  return s;  // valid in thread 0 only
}

}  // namespace fvx
