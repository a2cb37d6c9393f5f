// Semi-planar YUV 4:2:0 (NV12 / NV21) to RGB, BGR, RGBA or BGRA on gfx950.
//
// Each thread converts a 4 x 2 pixel block: 4 luma bytes from each of two
// rows and the 2 chroma pairs they share.  BT.601 limited range in 20-bit
// fixed point:
//   Y' = max(Y - 16, 0), U' = U - 128, V' = V - 128
//   R = (1220542 Y' + 1673527 V' + 2^19) >> 20
//   G = (1220542 Y' - 409993 U' - 852492 V' + 2^19) >> 20
//   B = (1220542 Y' + 2116026 U' + 2^19) >> 20,  saturated to u8.
// Reference: src/conversions/yuv420_to_rgb_neon.h:17-259,
// src/conversions/yuv42x_coefficients.h:95-105,
// src/conversions/yuv420sp_to_rgb_neon.cpp:100-160.

#include "kcv_internal.h"

namespace kcv {
namespace {

constexpr int kThreads = 256;
constexpr int kBlockW = 4;

__device__ __forceinline__ uint32_t sat_u8(int v) {
  return static_cast<uint32_t>(min(max(v, 0), 255));
}

template <int OC>
__global__ __launch_bounds__(kThreads) void nv12_kernel(YuvArgs a) {
  const int x0 = (blockIdx.x * kThreads + threadIdx.x) * kBlockW;
  if (x0 >= a.width) return;
  for (int yp = blockIdx.y; 2 * yp < a.height; yp += gridDim.y) {
    const uint8_t *uvrow = a.uv + static_cast<size_t>(yp) * a.uv_stride;
    int ru[2], gu[2], bu[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int x = min(x0 + 2 * p, a.width - 1) & ~1;
      const int u = uvrow[x + (a.vu ? 1 : 0)] - 128;
      const int v = uvrow[x + (a.vu ? 0 : 1)] - 128;
      ru[p] = 1673527 * v + (1 << 19);
      gu[p] = -409993 * u - 852492 * v + (1 << 19);
      bu[p] = 2116026 * u + (1 << 19);
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int y = 2 * yp + r;
      if (y >= a.height) break;
      const uint8_t *yrow = a.y + static_cast<size_t>(y) * a.y_stride;
      uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride;
#pragma unroll
      for (int k = 0; k < kBlockW; ++k) {
        const int x = x0 + k;
        if (x >= a.width) break;
        const int yw = 1220542 * max(yrow[x] - 16, 0);
        const int p = k >> 1;
        uint32_t R = sat_u8((yw + ru[p]) >> 20);
        const uint32_t G = sat_u8((yw + gu[p]) >> 20);
        uint32_t B = sat_u8((yw + bu[p]) >> 20);
        if (a.bgr) {
          const uint32_t t = R;
          R = B;
          B = t;
        }
        uint8_t *o = out + static_cast<size_t>(x) * OC;
        o[0] = static_cast<uint8_t>(R);
        o[1] = static_cast<uint8_t>(G);
        o[2] = static_cast<uint8_t>(B);
        if (OC == 4) o[3] = 0xFF;
      }
    }
  }
}

}  // namespace

hipError_t launch_yuv_sp_to_rgb(const YuvArgs &a, hipStream_t stream) {
  if (a.width <= 0 || a.height <= 0) return hipSuccess;
  const int per_block = kThreads * kBlockW;
  const int pairs = (a.height + 1) / 2;
  const dim3 grid((a.width + per_block - 1) / per_block, min(pairs, 65535));
  if (a.alpha)
    nv12_kernel<4><<<grid, kThreads, 0, stream>>>(a);
  else
    nv12_kernel<3><<<grid, kThreads, 0, stream>>>(a);
  return hipGetLastError();
}

}  // namespace kcv
