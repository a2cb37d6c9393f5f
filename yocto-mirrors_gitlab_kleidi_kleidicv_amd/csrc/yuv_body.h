// Device body of semi-planar YUV 4:2:0 (NV12 / NV21) to RGB, BGR, RGBA or
// BGRA on gfx950 (per-image kernel in yuv.hip, batched kernel in batch.hip).
//
// Each thread converts an 8 x 2 pixel block: 8 luma bytes from each of two
// rows and the 4 chroma pairs they share, read as 8-byte loads, written as
// 8- or 16-byte stores.  BT.601 limited range in 20-bit fixed point:
//   Y' = max(Y - 16, 0), U' = U - 128, V' = V - 128
//   R = (1220542 Y' + 1673527 V' + 2^19) >> 20
//   G = (1220542 Y' - 409993 U' - 852492 V' + 2^19) >> 20
//   B = (1220542 Y' + 2116026 U' + 2^19) >> 20,  saturated to u8.
// Reference: src/conversions/yuv420_to_rgb_neon.h:17-259,
// src/conversions/yuv42x_coefficients.h:95-105,
// src/conversions/yuv420sp_to_rgb_neon.cpp:100-160.

#ifndef KCV_YUV_BODY_H
#define KCV_YUV_BODY_H

#include "kcv_internal.h"

namespace kcv {
namespace yuv {

constexpr int kThreads = 256;
constexpr int kBlockW = 8;

__device__ __forceinline__ uint32_t sat_u8(int v) {
  return static_cast<uint32_t>(min(max(v, 0), 255));
}

// Column block bx (kThreads * kBlockW pixels) of row pairs yp_begin,
// yp_begin + yp_step, ... below yp_end.
template <int OC>
__device__ __forceinline__ void nv12_rows(const YuvArgs &a, int bx, int yp_begin, int yp_end,
                                          int yp_step) {
  const int x0 = (bx * kThreads + threadIdx.x) * kBlockW;
  if (x0 >= a.width) return;
  const bool full = x0 + kBlockW <= a.width;
  const bool in_al = ((reinterpret_cast<uintptr_t>(a.y) | a.y_stride |
                       reinterpret_cast<uintptr_t>(a.uv) | a.uv_stride) & 7) == 0;
  const bool out_al = ((reinterpret_cast<uintptr_t>(a.dst) | a.dst_stride) &
                       (OC == 4 ? 15 : 7)) == 0;
  for (int yp = yp_begin; yp < yp_end && 2 * yp < a.height; yp += yp_step) {
    const uint8_t *uvrow = a.uv + static_cast<size_t>(yp) * a.uv_stride;
    const bool two = 2 * yp + 1 < a.height;
    // All three loads (chroma row and both luma rows) are issued before any
    // arithmetic, so their latencies overlap.
    uint32_t uv_lo = 0, uv_hi = 0, yw_lo[2] = {0, 0}, yw_hi[2] = {0, 0};
    if (full && in_al) {
      const uint8_t *y0row = a.y + static_cast<size_t>(2 * yp) * a.y_stride + x0;
      const uint2 q = *reinterpret_cast<const uint2 *>(uvrow + x0);
      const uint2 q0 = *reinterpret_cast<const uint2 *>(y0row);
      const uint2 q1 = two ? *reinterpret_cast<const uint2 *>(y0row + a.y_stride) : q0;
      uv_lo = q.x;
      uv_hi = q.y;
      yw_lo[0] = q0.x;
      yw_hi[0] = q0.y;
      yw_lo[1] = q1.x;
      yw_hi[1] = q1.y;
    } else {
      for (int k = 0; k < kBlockW; ++k) {
        const int x = min(x0 + k, ((a.width + 1) & ~1) - 1);
        const uint32_t v = uvrow[x];
        if (k < 4) uv_lo |= v << (8 * k); else uv_hi |= v << (8 * (k - 4));
      }
      for (int r = 0; r < (two ? 2 : 1); ++r) {
        const uint8_t *yrow = a.y + static_cast<size_t>(2 * yp + r) * a.y_stride;
        for (int k = 0; k < kBlockW && x0 + k < a.width; ++k) {
          const uint32_t v = yrow[x0 + k];
          if (k < 4) yw_lo[r] |= v << (8 * k); else yw_hi[r] |= v << (8 * (k - 4));
        }
      }
    }
    // chroma terms of the 4 pairs
    int ru[4], gu[4], bu[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const uint32_t w = p < 2 ? uv_lo : uv_hi;
      const int b0 = static_cast<int>((w >> (16 * (p & 1))) & 0xFF);
      const int b1 = static_cast<int>((w >> (16 * (p & 1) + 8)) & 0xFF);
      const int u = (a.vu ? b1 : b0) - 128;
      const int v = (a.vu ? b0 : b1) - 128;
      ru[p] = 1673527 * v + (1 << 19);
      gu[p] = -409993 * u - 852492 * v + (1 << 19);
      bu[p] = 2116026 * u + (1 << 19);
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int y = 2 * yp + r;
      if (y >= a.height) break;
      const uint32_t ylo = yw_lo[r], yhi = yw_hi[r];
      uint32_t px[kBlockW];  // packed output pixels (R, G, B, A bytes)
#pragma unroll
      for (int k = 0; k < kBlockW; ++k) {
        const uint32_t yv = ((k < 4 ? ylo : yhi) >> (8 * (k & 3))) & 0xFF;
        const int yw = 1220542 * max(static_cast<int>(yv) - 16, 0);
        const int p = k >> 1;
        uint32_t R = sat_u8((yw + ru[p]) >> 20);
        const uint32_t G = sat_u8((yw + gu[p]) >> 20);
        uint32_t B = sat_u8((yw + bu[p]) >> 20);
        if (a.bgr) {
          const uint32_t t = R;
          R = B;
          B = t;
        }
        px[k] = R | (G << 8) | (B << 16) | (OC == 4 ? 0xFF000000u : 0u);
      }
      uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride + static_cast<size_t>(x0) * OC;
      if (full && out_al) {
        if (OC == 4) {
          store_nt(reinterpret_cast<uint4 *>(out), make_uint4(px[0], px[1], px[2], px[3]));
          store_nt(reinterpret_cast<uint4 *>(out) + 1, make_uint4(px[4], px[5], px[6], px[7]));
        } else {
          // 8 RGB pixels = 24 bytes = three 8-byte words
          uint32_t w[6];
#pragma unroll
          for (int k = 0; k < 6; ++k) w[k] = 0;
#pragma unroll
          for (int k = 0; k < kBlockW; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const int byte = 3 * k + c;
              w[byte >> 2] |= ((px[k] >> (8 * c)) & 0xFF) << (8 * (byte & 3));
            }
          uint2 *o = reinterpret_cast<uint2 *>(out);
          store_nt(o, make_uint2(w[0], w[1]));
          store_nt(o + 1, make_uint2(w[2], w[3]));
          store_nt(o + 2, make_uint2(w[4], w[5]));
        }
      } else {
        for (int k = 0; k < kBlockW && x0 + k < a.width; ++k)
#pragma unroll
          for (int c = 0; c < OC; ++c) out[k * OC + c] = static_cast<uint8_t>(px[k] >> (8 * c));
      }
    }
  }
}

}  // namespace yuv
}  // namespace kcv

#endif  // KCV_YUV_BODY_H
