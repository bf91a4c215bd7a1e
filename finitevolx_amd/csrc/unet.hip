// Learned preconditioner: a bias-free convolutional U-Net (3x3 convs, 2x2
// average pooling, nearest upsampling, additive skips) with BF16 weights and
// activations and FP32 accumulation on CDNA4 matrix cores.  No reference
// counterpart; the CPU oracle is oracle/unet.py.
//
// Layout: activations channels-last (H, W, C) bf16, C a multiple of 16 except
// the 1-channel input / output handled by dedicated FP32 kernels.  Convs with
// Cin, Cout multiples of 16 are implicit GEMMs on
// v_mfma_f32_16x16x16_bf16: M = Cout, N = 16 pixels of a row, K = 16
// input channels of one tap; a workgroup (4 waves) owns a 64 x 4 pixel tile
// whose input (+1 halo, optionally read through nearest upsampling) is
// staged in LDS; weights are pre-packed into per-lane MFMA fragments.

#include <cmath>
#include <cstring>
#include <vector>

#include "fvx_common.hpp"

namespace fvx {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int kUTw = 64, kUTh = 4, kUThreads = 256;

// ---------------------------------------------------------------------------
// Weight packing: fp32 w[9][Co][Ci] -> bf16 fragments
//   wp[((tap * COB + cob) * CIB + cib) * 64 + lane][i] =
//     W(tap, o = cob*16 + lane%16, c = cib*16 + 4*(lane/16) + i)
// with W(tap, o, c) = w[tap][o][c], or for the data gradient
// W(tap, o, c) = w[8 - tap][c][o] (flipped taps, swapped channels).
// ---------------------------------------------------------------------------
__global__ void pack_weights_kernel(const float* __restrict__ w, uint16_t* __restrict__ wp,
                                    int Co, int Ci, int flip) {
  // Co, Ci are the packed (output, input) channel counts.
  const int COB = Co / 16, CIB = Ci / 16;
  const int total = 9 * COB * CIB * 64 * 4;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    const int i = e & 3, lane = (e >> 2) & 63;
    int r = e >> 8;
    const int cib = r % CIB;
    r /= CIB;
    const int cob = r % COB;
    const int tap = r / COB;
    const int o = cob * 16 + (lane & 15), c = cib * 16 + 4 * (lane >> 4) + i;
    float v;
    if (!flip) v = w[((int64_t)tap * Co + o) * Ci + c];
    else v = w[((int64_t)(8 - tap) * Ci + c) * Co + o];  // source is w[9][Ci_src=Co][Co_src=Ci]
    wp[e] = f2bf(v);
  }
}

// ---------------------------------------------------------------------------
// 3x3 conv on matrix cores.  out = epi(sum_tap W x), epi: + skip, relu,
// * (maskt > 0).  UPS: x is (H/2, W/2) and read through nearest upsampling.
// ---------------------------------------------------------------------------
template <int COB, bool UPS>
__global__ __launch_bounds__(kUThreads) void conv_mfma_kernel(
    const uint16_t* __restrict__ x, int Cin, const bf16x4* __restrict__ wp,
    uint16_t* __restrict__ out, const uint16_t* __restrict__ skip,
    const uint16_t* __restrict__ maskt, int relu, int H, int W, int tiles_x, int ntiles,
    const int* __restrict__ guard) {
  if (guard && *guard) return;
  extern __shared__ uint16_t sx[];
  constexpr int Cout = COB * 16;
  const int CIB = Cin / 16;
  const int SW = kUTw + 2;
  const int t = xcd_swizzle(blockIdx.x, ntiles);
  const int i0 = (t % tiles_x) * kUTw, j0 = (t / tiles_x) * kUTh;
  const int tid = threadIdx.x;
  const int Hx = UPS ? H / 2 : H, Wx = UPS ? W / 2 : W;
  const int c4n = Cin / 4;
  const int nchunk = (kUTh + 2) * SW * c4n;
  for (int k = tid; k < nchunk; k += kUThreads) {
    const int pix = k / c4n, c4 = k % c4n;
    const int ry = pix / SW, rx = pix % SW;
    const int gy = j0 - 1 + ry, gx = i0 - 1 + rx;
    uint2 v = make_uint2(0u, 0u);
    if (gy >= 0 && gy < H && gx >= 0 && gx < W) {
      const int sy = UPS ? (gy >> 1) : gy, sxx = UPS ? (gx >> 1) : gx;
      v = *reinterpret_cast<const uint2*>(x + ((int64_t)sy * Wx + sxx) * Cin + 4 * c4);
    }
    *reinterpret_cast<uint2*>(sx + (int64_t)pix * Cin + 4 * c4) = v;
  }
  (void)Hx;
  __syncthreads();

  const int wave = tid >> 6, lane = tid & 63;
  const int lrow = lane & 15, lgrp = lane >> 4;
  f32x4 acc[COB][4];
#pragma unroll
  for (int a = 0; a < COB; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int tap = 0; tap < 9; ++tap) {
    const int ky = tap / 3, kx = tap % 3;
    for (int cib = 0; cib < CIB; ++cib) {
      bf16x4 bfrag[4];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        const int col = nb * 16 + lrow;
        const int64_t off = ((int64_t)(wave + ky) * SW + (col + kx)) * Cin + cib * 16 + 4 * lgrp;
        bfrag[nb] = *reinterpret_cast<const bf16x4*>(sx + off);
      }
#pragma unroll
      for (int cob = 0; cob < COB; ++cob) {
        const bf16x4 afrag = wp[((tap * COB + cob) * CIB + cib) * 64 + lane];
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
          acc[cob][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afrag, bfrag[nb], acc[cob][nb],
                                                                   0, 0, 0);
      }
    }
  }

  const int gy = j0 + wave;
  if (gy >= H) return;
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const int gx = i0 + nb * 16 + lrow;
    if (gx >= W) continue;
    const int64_t p = (int64_t)gy * W + gx;
#pragma unroll
    for (int cob = 0; cob < COB; ++cob) {
      const int c0 = cob * 16 + 4 * lgrp;
      uint16_t o4[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float v = acc[cob][nb][i];
        if (skip) v += bf2f(skip[p * Cout + c0 + i]);
        if (relu) v = fmaxf(v, 0.f);
        if (maskt) v = (bf2f(maskt[p * Cout + c0 + i]) > 0.f) ? v : 0.f;
        o4[i] = f2bf(v);
      }
      *reinterpret_cast<uint2*>(out + p * Cout + c0) =
          make_uint2((uint32_t)o4[0] | ((uint32_t)o4[1] << 16), (uint32_t)o4[2] | ((uint32_t)o4[3] << 16));
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient on matrix cores: dW[tap][co][ci] += sum_p gy[p][co] x[p+tap][ci]
// (x read through upsampling when UPS).  Each workgroup accumulates its own
// fp32 slot of ``part`` over a strided set of tiles; ``reduce`` sums slots.
// ---------------------------------------------------------------------------
template <bool UPS>
__global__ __launch_bounds__(kUThreads) void conv_bwdw_mfma_kernel(
    const uint16_t* __restrict__ x, int Cin, const uint16_t* __restrict__ gy, int Cout,
    float* __restrict__ part, int H, int W, int tiles_x, int ntiles) {
  extern __shared__ uint16_t smem[];
  const int SW = kUTw + 2;
  uint16_t* sx = smem;                                    // (TH+2) x SW x Cin
  uint16_t* sg = smem + (int64_t)(kUTh + 2) * SW * Cin;   // TH x TW x Cout
  const int COB = Cout / 16, CIB = Cin / 16;
  const int nblk = 9 * COB * CIB;
  float* mine = part + (int64_t)blockIdx.x * 9 * Cout * Cin;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lrow = lane & 15, lgrp = lane >> 4;
  const int Wx = UPS ? W / 2 : W;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int i0 = (t % tiles_x) * kUTw, j0 = (t / tiles_x) * kUTh;
    __syncthreads();
    const int c4x = Cin / 4;
    for (int k = tid; k < (kUTh + 2) * SW * c4x; k += kUThreads) {
      const int pix = k / c4x, c4 = k % c4x;
      const int ry = pix / SW, rx = pix % SW;
      const int gyy = j0 - 1 + ry, gxx = i0 - 1 + rx;
      uint2 v = make_uint2(0u, 0u);
      if (gyy >= 0 && gyy < H && gxx >= 0 && gxx < W) {
        const int sy = UPS ? (gyy >> 1) : gyy, sxx = UPS ? (gxx >> 1) : gxx;
        v = *reinterpret_cast<const uint2*>(x + ((int64_t)sy * Wx + sxx) * Cin + 4 * c4);
      }
      *reinterpret_cast<uint2*>(sx + (int64_t)pix * Cin + 4 * c4) = v;
    }
    const int c4g = Cout / 4;
    for (int k = tid; k < kUTh * kUTw * c4g; k += kUThreads) {
      const int pix = k / c4g, c4 = k % c4g;
      const int ry = pix / kUTw, rx = pix % kUTw;
      const int gyy = j0 + ry, gxx = i0 + rx;
      uint2 v = make_uint2(0u, 0u);
      if (gyy < H && gxx < W)
        v = *reinterpret_cast<const uint2*>(gy + ((int64_t)gyy * W + gxx) * Cout + 4 * c4);
      *reinterpret_cast<uint2*>(sg + (int64_t)pix * Cout + 4 * c4) = v;
    }
    __syncthreads();
    for (int b = wave; b < nblk; b += 4) {
      const int cib = b % CIB, cob = (b / CIB) % COB, tap = b / (CIB * COB);
      const int ky = tap / 3, kx = tap % 3;
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int kb = 0; kb < (kUTh * kUTw) / 16; ++kb) {
        bf16x4 a, bb;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int pix = kb * 16 + 4 * lgrp + i;
          const int ry = pix / kUTw, rx = pix % kUTw;
          a[i] = (short)sg[(int64_t)pix * Cout + cob * 16 + lrow];
          bb[i] = (short)sx[((int64_t)(ry + ky) * SW + rx + kx) * Cin + cib * 16 + lrow];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int co = cob * 16 + 4 * lgrp + i, ci = cib * 16 + lrow;
        mine[((int64_t)tap * Cout + co) * Cin + ci] += acc[i];
      }
    }
  }
}

__global__ void reduce_slots_kernel(const float* __restrict__ part, int nslot, int n,
                                    float* __restrict__ out, int accumulate) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int g = 0; g < nslot; ++g) s += part[(int64_t)g * n + e];
    out[e] = accumulate ? out[e] + s : s;
  }
}

// ---------------------------------------------------------------------------
// 1-channel input / output layers (FP32 arithmetic).
// ---------------------------------------------------------------------------
// out[p][c] = relu(sum_tap w[tap][c] r[p+tap]),  r fp32 (H, W), w fp32 [9][C]
__global__ void conv_first_kernel(const float* __restrict__ r, const float* __restrict__ w, int C,
                                  uint16_t* __restrict__ out, int H, int W,
                                  const int* __restrict__ guard) {
  if (guard && *guard) return;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= W || j >= H) return;
  float nb[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int y = j + t / 3 - 1, x = i + t % 3 - 1;
    nb[t] = (y >= 0 && y < H && x >= 0 && x < W) ? r[(int64_t)y * W + x] : 0.f;
  }
  const int64_t p = (int64_t)j * W + i;
  // 8 channels per 16-byte store (C is a multiple of 16).
  for (int c8 = 0; c8 < C; c8 += 8) {
    uint32_t pk[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      float s0 = 0.f, s1 = 0.f;
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        s0 += w[t * C + c8 + 2 * h] * nb[t];
        s1 += w[t * C + c8 + 2 * h + 1] * nb[t];
      }
      pk[h] = (uint32_t)f2bf(fmaxf(s0, 0.f)) | ((uint32_t)f2bf(fmaxf(s1, 0.f)) << 16);
    }
    *reinterpret_cast<uint4*>(out + p * C + c8) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
  }
}

// y[p] = scale * sum_tap sum_c w[tap][c] d[p+tap][c]; optional mask multiply.
// Each pixel's channels are read 8 at a time (16-byte loads).
template <typename TO>
__global__ void conv_last_kernel(const uint16_t* __restrict__ d, const float* __restrict__ w,
                                 int C, TO* __restrict__ y, int64_t ld_y, int H, int W,
                                 const float* __restrict__ scale, const uint8_t* __restrict__ mask,
                                 const int* __restrict__ guard) {
  if (guard && *guard) return;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= W || j >= H) return;
  float s = 0.f;
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int yy = j + t / 3 - 1, xx = i + t % 3 - 1;
    if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
    const uint16_t* dp = d + ((int64_t)yy * W + xx) * C;
    for (int c8 = 0; c8 < C; c8 += 8) {
      const uint4 v = *reinterpret_cast<const uint4*>(dp + c8);
      const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        s += w[t * C + c8 + 2 * h] * bf2f((uint16_t)(u[h] & 0xffffu));
        s += w[t * C + c8 + 2 * h + 1] * bf2f((uint16_t)(u[h] >> 16));
      }
    }
  }
  if (scale) s *= *scale;
  if (mask && !mask[(int64_t)j * W + i]) s = 0.f;
  y[(int64_t)j * ld_y + i] = (TO)s;
}

// g_d[q][c] = (d[q][c] > 0) * sum_tap w[tap][c] gy[q - tap]
__global__ void conv_last_bwd_data_kernel(const float* __restrict__ gy, const float* __restrict__ w,
                                          int C, const uint16_t* __restrict__ d,
                                          uint16_t* __restrict__ gd, int H, int W) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= W || j >= H) return;
  float g[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int y = j - (t / 3 - 1), x = i - (t % 3 - 1);
    g[t] = (y >= 0 && y < H && x >= 0 && x < W) ? gy[(int64_t)y * W + x] : 0.f;
  }
  const int64_t q = (int64_t)j * W + i;
  for (int c = 0; c < C; ++c) {
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < 9; ++t) s += w[t * C + c] * g[t];
    gd[q * C + c] = f2bf(bf2f(d[q * C + c]) > 0.f ? s : 0.f);
  }
}

// Per-block partial sums of dW[tap][c] = sum_p a[p] * b[p + tap][c] (a fp32
// single channel, b bf16 C channels) or, with swap, sum_p a[p][c] b[p + tap]
// (a bf16 C channels, b fp32 single channel).  part: [nblk][9*C].
__global__ __launch_bounds__(256) void edge_bwdw_kernel(const float* __restrict__ s1,
                                                         const uint16_t* __restrict__ sc, int C,
                                                         int swap, float* __restrict__ part,
                                                         int H, int W) {
  __shared__ float red[256];
  const int tid = threadIdx.x;
  const int64_t N = (int64_t)H * W;
  const int64_t per = (N + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = (int64_t)blockIdx.x * per, p1 = min(N, p0 + per);
  for (int e = 0; e < 9 * C; ++e) {
    const int t = e / C, c = e % C;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    float acc = 0.f;
    for (int64_t p = p0 + tid; p < p1; p += 256) {
      const int j = (int)(p / W), i = (int)(p % W);
      const int y = j + dy, x = i + dx;
      if (y < 0 || y >= H || x < 0 || x >= W) continue;
      const int64_t q = (int64_t)y * W + x;
      acc += swap ? bf2f(sc[p * C + c]) * s1[q] : s1[p] * bf2f(sc[q * C + c]);
    }
    red[tid] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    if (tid == 0) part[(int64_t)blockIdx.x * 9 * C + e] = red[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Pooling and gradient combination (bf16 HWC).
// ---------------------------------------------------------------------------
__global__ void pool2_kernel(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int H,
                             int W, int C, const int* __restrict__ guard) {  // H, W: output dims
  if (guard && *guard) return;
  const int64_t n = (int64_t)H * W * C;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t p = e / C;
    const int I = (int)(p % W), J = (int)(p / W);
    const int64_t W2 = 2 * W;
    const float s = bf2f(in[((2 * J) * W2 + 2 * I) * C + c]) + bf2f(in[((2 * J) * W2 + 2 * I + 1) * C + c]) +
                    bf2f(in[((2 * J + 1) * W2 + 2 * I) * C + c]) +
                    bf2f(in[((2 * J + 1) * W2 + 2 * I + 1) * C + c]);
    out[e] = f2bf(0.25f * s);
  }
}

// out[J][I][c] = (h[J][I][c] > 0) * sum_{a,b} g[2J+a][2I+b][c]   (up2 backward + relu mask)
__global__ void upbwd_mask_kernel(const uint16_t* __restrict__ g, const uint16_t* __restrict__ h,
                                  uint16_t* __restrict__ out, int H, int W, int C) {
  const int64_t n = (int64_t)H * W * C;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t p = e / C;
    const int I = (int)(p % W), J = (int)(p / W);
    const int64_t W2 = 2 * W;
    const float s = bf2f(g[((2 * J) * W2 + 2 * I) * C + c]) + bf2f(g[((2 * J) * W2 + 2 * I + 1) * C + c]) +
                    bf2f(g[((2 * J + 1) * W2 + 2 * I) * C + c]) +
                    bf2f(g[((2 * J + 1) * W2 + 2 * I + 1) * C + c]);
    out[e] = f2bf(bf2f(h[e]) > 0.f ? s : 0.f);
  }
}

// out[j][i][c] = (h[j][i][c] > 0) * (skip[j][i][c] + 0.25 gc[j/2][i/2][c])  (pool backward)
__global__ void poolbwd_add_mask_kernel(const uint16_t* __restrict__ skip,
                                        const uint16_t* __restrict__ gc,
                                        const uint16_t* __restrict__ h, uint16_t* __restrict__ out,
                                        int H, int W, int C) {
  const int64_t n = (int64_t)H * W * C;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t p = e / C;
    const int i = (int)(p % W), j = (int)(p / W);
    const float v = (skip ? bf2f(skip[e]) : 0.f) +
                    0.25f * bf2f(gc[((int64_t)(j / 2) * (W / 2) + i / 2) * C + c]);
    out[e] = f2bf(bf2f(h[e]) > 0.f ? v : 0.f);
  }
}

// ---------------------------------------------------------------------------
// Input normalisation, loss, Adam.
// ---------------------------------------------------------------------------
// s = rms of the residual (FP64 accumulate): per-block partials, then a
// one-workgroup finish (kernel boundary in between, no device fences):
// scal[0] = 1/s (input scale), scal[1] = s * out_scale (output scale).
template <typename T>
__global__ __launch_bounds__(256) void rms_kernel(const T* __restrict__ r, int64_t ld, int H, int W,
                                                   double* __restrict__ part,
                                                   const int* __restrict__ guard) {
  if (guard && *guard) return;
  __shared__ double red[256];
  const int tid = threadIdx.x;
  double acc = 0.0;
  const int64_t N = (int64_t)H * W;
  for (int64_t p = blockIdx.x * 256 + tid; p < N; p += (int64_t)gridDim.x * 256) {
    const double v = (double)r[(p / W) * ld + p % W];
    acc += v * v;
  }
  red[tid] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void rms_finish_kernel(const double* __restrict__ part, int nb,
                                                          int64_t N, float* __restrict__ scal,
                                                          double out_scale,
                                                          const int* __restrict__ guard) {
  if (guard && *guard) return;
  __shared__ double red[256];
  const int tid = threadIdx.x;
  double acc = 0.0;
  for (int b = tid; b < nb; b += 256) acc += part[b];
  red[tid] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) {
    const double rms = sqrt(red[0] / (double)N);
    scal[0] = rms > 0.0 ? (float)(1.0 / rms) : 0.f;
    scal[1] = (float)(rms * out_scale);
  }
}

template <typename T>
__global__ void scale_in_kernel(const T* __restrict__ r, int64_t ld, int H, int W,
                                const float* __restrict__ scal, float* __restrict__ out,
                                const int* __restrict__ guard) {
  if (guard && *guard) return;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i < W && j < H) out[(int64_t)j * W + i] = (float)r[(int64_t)j * ld + i] * scal[0];
}

// gy = 2 (y - e) / n  (loss = mean (y - e)^2), loss partial sums in part.
__global__ __launch_bounds__(256) void mse_grad_kernel(const float* __restrict__ y,
                                                       const float* __restrict__ e,
                                                       float* __restrict__ gy, int64_t n,
                                                       double* __restrict__ loss_acc) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  double acc = 0.0;
  for (int64_t k = blockIdx.x * 256 + tid; k < n; k += (int64_t)gridDim.x * 256) {
    const float d = y[k] - e[k];
    gy[k] = 2.f * d / (float)n;
    acc += (double)d * d;
  }
  red[tid] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) atomicAdd(loss_acc, red[0] / (double)n);
}

__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2, float gscale) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const float gk = g[k] * gscale;
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
    g[k] = 0.f;
  }
}

// ---------------------------------------------------------------------------
// Network plan
// ---------------------------------------------------------------------------
struct Layer {
  int cout, cin;
  size_t off;  // offset (floats) into the master weight buffer
};

struct UNet {
  int L;                    // levels
  int ch[8];
  int H, W;                 // spatial size the buffers are sized for
  std::vector<Layer> layers;  // order: e0a e0b e1a e1b ... u_{L-2} d_{L-2} ... u0 d0 out
  size_t nweights = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // fp32 master / grad / Adam
  uint16_t* wpk = nullptr;   // packed bf16 forward fragments (per layer offsets)
  uint16_t* wpkT = nullptr;  // packed flipped/transposed fragments (data gradient)
  std::vector<size_t> pk_off;
  // activations per level: x (pooled input, l > 0), h1e, e (=h2e), h1d, d
  std::vector<uint16_t*> xin, h1e, ee, h1d, dd, gA, gB, gC;
  float *rin = nullptr, *yout = nullptr, *gy = nullptr, *eref = nullptr;
  float* part = nullptr;
  size_t part_elems = 0;
  double* rms_part = nullptr;
  float* scal = nullptr;
  double* loss = nullptr;
  const int* guard = nullptr;  // PCG convergence flag: kernels exit when set
  int adam_t = 0;
  std::vector<void*> allocs;
};

static int layer_index(const UNet& N, char kind, int l) {
  // e: 2l (a) / 2l+1 (b); u/d: 2L + 2*(L-2-l) (+1 for d); out: last
  if (kind == 'a') return 2 * l;
  if (kind == 'b') return 2 * l + 1;
  if (kind == 'u') return 2 * N.L + 2 * (N.L - 2 - l);
  if (kind == 'd') return 2 * N.L + 2 * (N.L - 2 - l) + 1;
  return (int)N.layers.size() - 1;
}

template <typename T>
static int dev_alloc(UNet& N, T** p, size_t n) {
  void* q = nullptr;
  FVX_HIP(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
  N.allocs.push_back(q);
  *p = (T*)q;
  return kOk;
}

static void conv_launch(const UNet& N, int li, bool transposed, const uint16_t* x, uint16_t* out,
                        const uint16_t* skip, const uint16_t* maskt, bool relu, bool ups, int H,
                        int W, hipStream_t s) {
  const Layer& Ly = N.layers[li];
  const int cout = transposed ? Ly.cin : Ly.cout, cin = transposed ? Ly.cout : Ly.cin;
  const bf16x4* wp = (const bf16x4*)((transposed ? N.wpkT : N.wpk) + N.pk_off[li]);
  const int tx = cdiv(W, kUTw), ty = cdiv(H, kUTh);
  const size_t lds = (size_t)(kUTh + 2) * (kUTw + 2) * cin * sizeof(uint16_t);
#define FVX_CONV(COB, UP)                                                                     \
  hipLaunchKernelGGL((conv_mfma_kernel<COB, UP>), dim3(tx * ty), dim3(kUThreads), lds, s, x, cin, \
                     wp, out, skip, maskt, relu ? 1 : 0, H, W, tx, tx * ty, N.guard)
  const int cob = cout / 16;
  if (ups) {
    if (cob == 1) FVX_CONV(1, true);
    else if (cob == 2) FVX_CONV(2, true);
    else if (cob == 3) FVX_CONV(3, true);
    else FVX_CONV(4, true);
  } else {
    if (cob == 1) FVX_CONV(1, false);
    else if (cob == 2) FVX_CONV(2, false);
    else if (cob == 3) FVX_CONV(3, false);
    else FVX_CONV(4, false);
  }
#undef FVX_CONV
}

static dim3 g2(int H, int W) { return dim3(cdiv(W, 64), cdiv(H, 4)); }
static dim3 g1(int64_t n) { return dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)); }

// Forward pass on N.rin (already scaled); result in N.yout (fp32), unscaled.
// With last == false the output conv is left to the caller.
static int unet_forward(UNet& N, int H, int W, hipStream_t s, bool last = true) {
  const int L = N.L;
  // encoder
  for (int l = 0; l < L; ++l) {
    const int h = H >> l, w = W >> l;
    if (l == 0) {
      const Layer& a = N.layers[layer_index(N, 'a', 0)];
      hipLaunchKernelGGL(conv_first_kernel, g2(h, w), dim3(64, 4), 0, s, N.rin, N.w + a.off,
                         N.ch[0], N.h1e[0], h, w, N.guard);
    } else {
      hipLaunchKernelGGL(pool2_kernel, g1((int64_t)h * w * N.ch[l - 1]), dim3(256), 0, s,
                         N.ee[l - 1], N.xin[l], h, w, N.ch[l - 1], N.guard);
      conv_launch(N, layer_index(N, 'a', l), false, N.xin[l], N.h1e[l], nullptr, nullptr, true,
                  false, h, w, s);
    }
    conv_launch(N, layer_index(N, 'b', l), false, N.h1e[l], N.ee[l], nullptr, nullptr, true, false,
                h, w, s);
  }
  // decoder
  const uint16_t* dprev = N.ee[L - 1];
  for (int l = L - 2; l >= 0; --l) {
    const int h = H >> l, w = W >> l;
    conv_launch(N, layer_index(N, 'u', l), false, dprev, N.h1d[l], N.ee[l], nullptr, true, true, h,
                w, s);
    conv_launch(N, layer_index(N, 'd', l), false, N.h1d[l], N.dd[l], nullptr, nullptr, true, false,
                h, w, s);
    dprev = N.dd[l];
  }
  if (last) {
    const Layer& o = N.layers.back();
    hipLaunchKernelGGL((conv_last_kernel<float>), g2(H, W), dim3(64, 4), 0, s, dprev, N.w + o.off,
                       N.ch[0], N.yout, (int64_t)W, H, W, nullptr, nullptr, N.guard);
  }
  FVX_LAUNCH_CHECK();
  return kOk;
}

static int bwdw_launch(UNet& N, int li, const uint16_t* x, const uint16_t* gyt, bool ups, int H,
                       int W, hipStream_t s) {
  const Layer& Ly = N.layers[li];
  const int tx = cdiv(W, kUTw), ty = cdiv(H, kUTh), nt = tx * ty;
  const int G = std::min(nt, 64);
  const size_t n = (size_t)9 * Ly.cout * Ly.cin;
  FVX_CHECK_ARG((size_t)G * n <= N.part_elems, "gradient scratch too small");
  FVX_HIP(hipMemsetAsync(N.part, 0, (size_t)G * n * sizeof(float), s));
  const size_t lds = ((size_t)(kUTh + 2) * (kUTw + 2) * Ly.cin + (size_t)kUTh * kUTw * Ly.cout) *
                     sizeof(uint16_t);
  if (ups)
    hipLaunchKernelGGL(conv_bwdw_mfma_kernel<true>, dim3(G), dim3(kUThreads), lds, s, x, Ly.cin,
                       gyt, Ly.cout, N.part, H, W, tx, nt);
  else
    hipLaunchKernelGGL(conv_bwdw_mfma_kernel<false>, dim3(G), dim3(kUThreads), lds, s, x, Ly.cin,
                       gyt, Ly.cout, N.part, H, W, tx, nt);
  hipLaunchKernelGGL(reduce_slots_kernel, g1((int64_t)n), dim3(256), 0, s, N.part, G, (int)n,
                     N.g + Ly.off, 1);
  FVX_LAUNCH_CHECK();
  return kOk;
}

static int edge_bwdw(UNet& N, int li, const float* s1, const uint16_t* sc, int C, bool swap,
                     int H, int W, hipStream_t s) {
  const Layer& Ly = N.layers[li];
  const int G = 64;
  const size_t n = (size_t)9 * C;
  hipLaunchKernelGGL(edge_bwdw_kernel, dim3(G), dim3(256), 0, s, s1, sc, C, swap ? 1 : 0, N.part,
                     H, W);
  hipLaunchKernelGGL(reduce_slots_kernel, g1((int64_t)n), dim3(256), 0, s, N.part, G, (int)n,
                     N.g + Ly.off, 1);
  FVX_LAUNCH_CHECK();
  return kOk;
}

// Backward from N.gy (dLoss/dy, fp32) to weight gradients accumulated in N.g.
static int unet_backward(UNet& N, int H, int W, hipStream_t s) {
  const int L = N.L;
  // output layer
  const Layer& o = N.layers.back();
  const uint16_t* d0 = (L > 1) ? N.dd[0] : N.ee[0];
  int rc = edge_bwdw(N, (int)N.layers.size() - 1, N.gy, d0, N.ch[0], false, H, W, s);
  if (rc) return rc;
  hipLaunchKernelGGL(conv_last_bwd_data_kernel, g2(H, W), dim3(64, 4), 0, s, N.gy, N.w + o.off,
                     N.ch[0], d0, N.gA[0], H, W);
  // decoder levels 0 .. L-2: gA[l] holds the masked gradient into d_l's a2
  for (int l = 0; l <= L - 2; ++l) {
    const int h = H >> l, w = W >> l;
    rc = bwdw_launch(N, layer_index(N, 'd', l), N.h1d[l], N.gA[l], false, h, w, s);
    if (rc) return rc;
    // ga1 = conv^T(ga2) * (h1d > 0) -> gB[l]  (also the skip gradient into e_l)
    conv_launch(N, layer_index(N, 'd', l), true, N.gA[l], N.gB[l], nullptr, N.h1d[l], false, false,
                h, w, s);
    const uint16_t* dnext = (l + 1 == L - 1) ? N.ee[L - 1] : N.dd[l + 1];
    rc = bwdw_launch(N, layer_index(N, 'u', l), dnext, N.gB[l], true, h, w, s);
    if (rc) return rc;
    // gu (fine, C_{l+1} channels) -> gC[l]; up2 backward + relu mask of d_{l+1} -> gA[l+1]
    conv_launch(N, layer_index(N, 'u', l), true, N.gB[l], N.gC[l], nullptr, nullptr, false, false,
                h, w, s);
    const int hc = H >> (l + 1), wc = W >> (l + 1), c = N.ch[l + 1];
    hipLaunchKernelGGL(upbwd_mask_kernel, g1((int64_t)hc * wc * c), dim3(256), 0, s, N.gC[l],
                       dnext, N.gA[l + 1], hc, wc, c);
  }
  // encoder levels L-1 .. 0: gA[l] holds the masked gradient into e_l's a2
  for (int l = L - 1; l >= 0; --l) {
    const int h = H >> l, w = W >> l;
    rc = bwdw_launch(N, layer_index(N, 'b', l), N.h1e[l], N.gA[l], false, h, w, s);
    if (rc) return rc;
    // ga1 = conv^T(ga2) * (h1e > 0) -> gC[l]
    conv_launch(N, layer_index(N, 'b', l), true, N.gA[l], N.gC[l], nullptr, N.h1e[l], false, false,
                h, w, s);
    if (l == 0) {
      rc = edge_bwdw(N, layer_index(N, 'a', 0), N.rin, N.gC[0], N.ch[0], true, h, w, s);
      if (rc) return rc;
      break;
    }
    rc = bwdw_launch(N, layer_index(N, 'a', l), N.xin[l], N.gC[l], false, h, w, s);
    if (rc) return rc;
    // gx = conv^T(ga1) (C_{l-1} channels at level l) -> gA[l] (free now)
    conv_launch(N, layer_index(N, 'a', l), true, N.gC[l], N.gA[l], nullptr, nullptr, false, false,
                h, w, s);
    // gradient into e_{l-1}: (skip grad gB[l-1] + pool^T gx) * (e_{l-1} > 0) -> gA[l-1]
    const int hf = H >> (l - 1), wf = W >> (l - 1), c = N.ch[l - 1];
    hipLaunchKernelGGL(poolbwd_add_mask_kernel, g1((int64_t)hf * wf * c), dim3(256), 0, s,
                       (l - 1 <= L - 2) ? N.gB[l - 1] : nullptr, N.gA[l], N.ee[l - 1],
                       N.gA[l - 1], hf, wf, c);
  }
  FVX_LAUNCH_CHECK();
  return kOk;
}

static int repack(UNet& N, hipStream_t s) {
  for (size_t li = 0; li < N.layers.size(); ++li) {
    const Layer& Ly = N.layers[li];
    if (Ly.cin % 16 || Ly.cout % 16) continue;  // edge layers use fp32 weights directly
    hipLaunchKernelGGL(pack_weights_kernel, dim3(64), dim3(256), 0, s, N.w + Ly.off,
                       N.wpk + N.pk_off[li], Ly.cout, Ly.cin, 0);
    hipLaunchKernelGGL(pack_weights_kernel, dim3(64), dim3(256), 0, s, N.w + Ly.off,
                       N.wpkT + N.pk_off[li], Ly.cin, Ly.cout, 1);
  }
  FVX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace fvx

extern "C" {

// channels: nlev entries (each a multiple of 16, <= 64); H, W: the largest
// spatial size the plan will run (divisible by 2^(nlev-1)).
int fvx_unet_create(int nlev, const int* channels, int H, int W, void** out) {
  using namespace fvx;
  FVX_CHECK_ARG(out && channels, "null argument");
  FVX_CHECK_ARG(nlev >= 1 && nlev <= 8, "1..8 levels");
  for (int l = 0; l < nlev; ++l)
    FVX_CHECK_ARG(channels[l] % 16 == 0 && channels[l] >= 16 && channels[l] <= 64,
                  "channels must be multiples of 16 in [16, 64]");
  FVX_CHECK_ARG(H % (1 << (nlev - 1)) == 0 && W % (1 << (nlev - 1)) == 0,
                "H, W must be divisible by 2^(nlev-1)");
  UNet* N = new UNet();
  N->L = nlev;
  for (int l = 0; l < nlev; ++l) N->ch[l] = channels[l];
  N->H = H;
  N->W = W;
  size_t off = 0, pk = 0;
  auto add = [&](int co, int ci) {
    N->layers.push_back({co, ci, off});
    N->pk_off.push_back(pk);
    off += (size_t)9 * co * ci;
    if (co % 16 == 0 && ci % 16 == 0) pk += (size_t)9 * co * ci;
  };
  int cin = 1;
  for (int l = 0; l < nlev; ++l) {
    add(channels[l], cin);
    add(channels[l], channels[l]);
    cin = channels[l];
  }
  for (int l = nlev - 2; l >= 0; --l) {
    add(channels[l], channels[l + 1]);
    add(channels[l], channels[l]);
  }
  add(1, channels[0]);
  N->nweights = off;
  int rc = 0;
  rc |= dev_alloc(*N, &N->w, off);
  rc |= dev_alloc(*N, &N->g, off);
  rc |= dev_alloc(*N, &N->m, off);
  rc |= dev_alloc(*N, &N->v, off);
  rc |= dev_alloc(*N, &N->wpk, pk);
  rc |= dev_alloc(*N, &N->wpkT, pk);
  N->xin.assign(nlev, nullptr);
  N->h1e.assign(nlev, nullptr);
  N->ee.assign(nlev, nullptr);
  N->h1d.assign(nlev, nullptr);
  N->dd.assign(nlev, nullptr);
  N->gA.assign(nlev, nullptr);
  N->gB.assign(nlev, nullptr);
  N->gC.assign(nlev, nullptr);
  for (int l = 0; l < nlev && !rc; ++l) {
    const size_t px = (size_t)(H >> l) * (W >> l);
    const int cmax = std::max(channels[l], l + 1 < nlev ? channels[l + 1] : 0);
    if (l > 0) rc |= dev_alloc(*N, &N->xin[l], px * channels[l - 1]);
    rc |= dev_alloc(*N, &N->h1e[l], px * channels[l]);
    rc |= dev_alloc(*N, &N->ee[l], px * channels[l]);
    rc |= dev_alloc(*N, &N->h1d[l], px * channels[l]);
    rc |= dev_alloc(*N, &N->dd[l], px * channels[l]);
    const int cg = std::max(cmax, l > 0 ? channels[l - 1] : 0);
    rc |= dev_alloc(*N, &N->gA[l], px * cg);
    rc |= dev_alloc(*N, &N->gB[l], px * cg);
    rc |= dev_alloc(*N, &N->gC[l], px * cg);
  }
  const size_t npx = (size_t)H * W;
  rc |= dev_alloc(*N, &N->rin, npx);
  rc |= dev_alloc(*N, &N->yout, npx);
  rc |= dev_alloc(*N, &N->gy, npx);
  rc |= dev_alloc(*N, &N->eref, npx);
  N->part_elems = (size_t)64 * 9 * 64 * 64;
  rc |= dev_alloc(*N, &N->part, N->part_elems);
  rc |= dev_alloc(*N, &N->rms_part, 4096);
  rc |= dev_alloc(*N, &N->scal, 4);
  rc |= dev_alloc(*N, &N->loss, 1);
  if (!rc) {
    const int big = 160 * 1024;
    (void)hipFuncSetAttribute((const void*)conv_bwdw_mfma_kernel<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, big);
    (void)hipFuncSetAttribute((const void*)conv_bwdw_mfma_kernel<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, big);
    hipError_t e = hipMemset(N->g, 0, off * sizeof(float));
    if (e == hipSuccess) e = hipMemset(N->m, 0, off * sizeof(float));
    if (e == hipSuccess) e = hipMemset(N->v, 0, off * sizeof(float));
    if (e != hipSuccess) rc = kHipError;
  }
  if (rc) {
    for (void* p : N->allocs) (void)hipFree(p);
    delete N;
    if (rc != kInvalidArgument) set_error("allocation failed for the U-Net plan");
    return rc == kInvalidArgument ? rc : kHipError;
  }
  *out = N;
  return kOk;
}

int fvx_unet_destroy(void* net) {
  using namespace fvx;
  if (!net) return kOk;
  UNet* N = (UNet*)net;
  (void)hipDeviceSynchronize();
  for (void* p : N->allocs) (void)hipFree(p);
  delete N;
  return kOk;
}

// Number of fp32 weights (all layers, layout [9][Cout][Cin] per layer in the
// order e0a e0b .. e{L-1}b, u{L-2} d{L-2} .. u0 d0, out).
int64_t fvx_unet_num_weights(void* net) { return net ? (int64_t)((fvx::UNet*)net)->nweights : -1; }

// Copy fp32 weights in (device pointer) or out; resets Adam state on "in".
int fvx_unet_weights(void* net, float* w, int direction, void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(net && w, "null argument");
  UNet* N = (UNet*)net;
  hipStream_t s = (hipStream_t)stream;
  if (direction == 0) {
    FVX_HIP(hipMemcpyAsync(N->w, w, N->nweights * sizeof(float), hipMemcpyDeviceToDevice, s));
    FVX_HIP(hipMemsetAsync(N->m, 0, N->nweights * sizeof(float), s));
    FVX_HIP(hipMemsetAsync(N->v, 0, N->nweights * sizeof(float), s));
    FVX_HIP(hipMemsetAsync(N->g, 0, N->nweights * sizeof(float), s));
    N->adam_t = 0;
    return repack(*N, s);
  }
  FVX_HIP(hipMemcpyAsync(w, N->w, N->nweights * sizeof(float), hipMemcpyDeviceToDevice, s));
  return kOk;
}

// Raw forward on an fp32 (H, W) input already in the network's units:
// y = net(x).  H, W <= the plan's size and divisible by 2^(L-1).
int fvx_unet_forward_raw(void* net, const float* x, float* y, int H, int W, void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(net && x && y, "null argument");
  UNet* N = (UNet*)net;
  FVX_CHECK_ARG(H <= N->H && W <= N->W && H * W <= N->H * N->W, "input larger than the plan");
  FVX_CHECK_ARG(H % (1 << (N->L - 1)) == 0 && W % (1 << (N->L - 1)) == 0, "size not divisible");
  hipStream_t s = (hipStream_t)stream;
  FVX_HIP(hipMemcpyAsync(N->rin, x, (size_t)H * W * sizeof(float), hipMemcpyDeviceToDevice, s));
  int rc = unet_forward(*N, H, W, s);
  if (rc) return rc;
  FVX_HIP(hipMemcpyAsync(y, N->yout, (size_t)H * W * sizeof(float), hipMemcpyDeviceToDevice, s));
  return kOk;
}

// Weight gradients of sum(gy * net(x)) accumulated into the plan's gradient
// buffer (read with fvx_unet_grads); for tests and training.
int fvx_unet_backward_raw(void* net, const float* x, const float* gy, int H, int W,
                          void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(net && x && gy, "null argument");
  UNet* N = (UNet*)net;
  FVX_CHECK_ARG(H <= N->H && W <= N->W, "input larger than the plan");
  hipStream_t s = (hipStream_t)stream;
  FVX_HIP(hipMemcpyAsync(N->rin, x, (size_t)H * W * sizeof(float), hipMemcpyDeviceToDevice, s));
  int rc = unet_forward(*N, H, W, s);
  if (rc) return rc;
  FVX_HIP(hipMemcpyAsync(N->gy, gy, (size_t)H * W * sizeof(float), hipMemcpyDeviceToDevice, s));
  return unet_backward(*N, H, W, s);
}

int fvx_unet_grads(void* net, float* g, int zero_after, void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(net && g, "null argument");
  UNet* N = (UNet*)net;
  hipStream_t s = (hipStream_t)stream;
  FVX_HIP(hipMemcpyAsync(g, N->g, N->nweights * sizeof(float), hipMemcpyDeviceToDevice, s));
  if (zero_after) FVX_HIP(hipMemsetAsync(N->g, 0, N->nweights * sizeof(float), s));
  return kOk;
}

// One training step on a batch of B pairs (x_b, e_b), each (H, W) fp32 in the
// network's units: loss = mean_b mean_p (net(x_b) - e_b)^2; Adam update.
// Returns the batch loss in *loss_out (synchronises the stream).
int fvx_unet_train_step(void* net, const float* xb, const float* eb, int B, int H, int W,
                        double lr, double* loss_out, void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(net && xb && eb && B >= 1, "bad arguments");
  UNet* N = (UNet*)net;
  FVX_CHECK_ARG(H <= N->H && W <= N->W, "batch item larger than the plan");
  hipStream_t s = (hipStream_t)stream;
  FVX_HIP(hipMemsetAsync(N->loss, 0, sizeof(double), s));
  const size_t npx = (size_t)H * W;
  for (int b = 0; b < B; ++b) {
    FVX_HIP(hipMemcpyAsync(N->rin, xb + b * npx, npx * sizeof(float), hipMemcpyDeviceToDevice, s));
    int rc = unet_forward(*N, H, W, s);
    if (rc) return rc;
    hipLaunchKernelGGL(mse_grad_kernel, g1((int64_t)npx), dim3(256), 0, s, N->yout, eb + b * npx,
                       N->gy, (int64_t)npx, N->loss);
    rc = unet_backward(*N, H, W, s);
    if (rc) return rc;
  }
  N->adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.f - std::pow(b1, (float)N->adam_t), bc2 = 1.f - std::pow(b2, (float)N->adam_t);
  hipLaunchKernelGGL(adam_kernel, g1((int64_t)N->nweights), dim3(256), 0, s, N->w, N->g, N->m,
                     N->v, (int64_t)N->nweights, (float)lr, b1, b2, 1e-8f, bc1, bc2, 1.0f / B);
  int rc = repack(*N, s);
  if (rc) return rc;
  double l = 0.0;
  FVX_HIP(hipMemcpyAsync(&l, N->loss, sizeof(double), hipMemcpyDeviceToHost, s));
  FVX_HIP(hipStreamSynchronize(s));
  if (loss_out) *loss_out = l / B;
  return kOk;
}

}  // extern "C"

// Preconditioner application for PCG: z = out_scale * rms(r) * net(r / rms(r)) * mask.
namespace fvx {
template <typename T>
int unet_apply(void* net, const T* r, int64_t ld_r, T* z, int64_t ld_z, int H, int W,
               double out_scale, const uint8_t* mask, hipStream_t s, const int* guard) {
  UNet* N = (UNet*)net;
  N->guard = guard;
  FVX_CHECK_ARG(H <= N->H && W <= N->W, "field larger than the network plan");
  FVX_CHECK_ARG(H % (1 << (N->L - 1)) == 0 && W % (1 << (N->L - 1)) == 0,
                "field size not divisible by 2^(levels-1)");
  const int nb = 256;
  hipLaunchKernelGGL(rms_kernel<T>, dim3(nb), dim3(256), 0, s, r, ld_r, H, W, N->rms_part,
                     N->guard);
  hipLaunchKernelGGL(rms_finish_kernel, dim3(1), dim3(256), 0, s, N->rms_part, nb,
                     (int64_t)H * W, N->scal, out_scale, N->guard);
  hipLaunchKernelGGL(scale_in_kernel<T>, g2(H, W), dim3(64, 4), 0, s, r, ld_r, H, W, N->scal,
                     N->rin, N->guard);
  int rc = unet_forward(*N, H, W, s, false);
  if (rc) return rc;
  // Output conv with the rescaling and mask fused.
  hipLaunchKernelGGL((conv_last_kernel<T>), g2(H, W), dim3(64, 4), 0, s,
                     (N->L > 1) ? N->dd[0] : N->ee[0], N->w + N->layers.back().off, N->ch[0], z,
                     ld_z, H, W, N->scal + 1, mask, N->guard);
  N->guard = nullptr;
  FVX_LAUNCH_CHECK();
  return kOk;
}
template int unet_apply<float>(void*, const float*, int64_t, float*, int64_t, int, int, double,
                               const uint8_t*, hipStream_t, const int*);
template int unet_apply<double>(void*, const double*, int64_t, double*, int64_t, int, int, double,
                                const uint8_t*, hipStream_t, const int*);
}  // namespace fvx

// Standalone 3x3 conv (bf16 HWC in/out, fp32 weights [9][Cout][Cin], Cin and
// Cout multiples of 16 <= 64), optionally reading x through nearest
// upsampling and applying ReLU.  ``scratch`` holds >= 9*Cout*Cin bf16.
extern "C" int fvx_conv3x3_bf16(const uint16_t* x, int Cin, const float* w, int Cout,
                                uint16_t* out, int H, int W, int ups, int relu, uint16_t* scratch,
                                void* stream) {
  using namespace fvx;
  FVX_CHECK_ARG(x && w && out && scratch, "null argument");
  FVX_CHECK_ARG(Cin % 16 == 0 && Cout % 16 == 0 && Cin <= 64 && Cout <= 64 && Cin > 0 && Cout > 0,
                "channels must be multiples of 16 in [16, 64]");
  FVX_CHECK_ARG(!ups || (H % 2 == 0 && W % 2 == 0), "upsampling needs even H, W");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(pack_weights_kernel, dim3(64), dim3(256), 0, s, w, scratch, Cout, Cin, 0);
  UNet N;
  N.layers.push_back({Cout, Cin, 0});
  N.pk_off.push_back(0);
  N.wpk = scratch;
  N.wpkT = scratch;
  conv_launch(N, 0, false, x, out, nullptr, nullptr, relu != 0, ups != 0, H, W, s);
  FVX_LAUNCH_CHECK();
  return kOk;
}

extern "C" int fvx_unet_apply_f32(void* net, const float* r, int64_t ld_r, float* z, int64_t ld_z,
                                  int H, int W, double out_scale, const uint8_t* mask,
                                  void* stream) {
  if (!net || !r || !z) {
    fvx::set_error("null argument");
    return fvx::kInvalidArgument;
  }
  return fvx::unet_apply<float>(net, r, ld_r, z, ld_z, H, W, out_scale, mask, (hipStream_t)stream,
                                nullptr);
}
