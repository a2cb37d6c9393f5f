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

// Chroma terms of 4 pixel pairs (bytes U0 V0 U1 V1 | U2 V2 U3 V3 of
// uv_lo | uv_hi, V first when `vu`), in the 20-bit fixed point above.
struct Chroma {
  int r[4], g[4], b[4];
};

__device__ __forceinline__ Chroma chroma4(uint32_t uv_lo, uint32_t uv_hi, bool vu) {
  Chroma t;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const uint32_t w = p < 2 ? uv_lo : uv_hi;
    const int b0 = static_cast<int>((w >> (16 * (p & 1))) & 0xFF);
    const int b1 = static_cast<int>((w >> (16 * (p & 1) + 8)) & 0xFF);
    const int u = (vu ? b1 : b0) - 128;
    const int v = (vu ? b0 : b1) - 128;
    t.r[p] = 1673527 * v + (1 << 19);
    t.g[p] = -409993 * u - 852492 * v + (1 << 19);
    t.b[p] = 2116026 * u + (1 << 19);
  }
  return t;
}

__device__ __forceinline__ uint32_t pack_px(uint32_t R, uint32_t G, uint32_t B, bool bgr,
                                            bool alpha) {
  if (bgr) {
    const uint32_t t = R;
    R = B;
    B = t;
  }
  return R | (G << 8) | (B << 16) | (alpha ? 0xFF000000u : 0u);
}

// 8 packed output pixels (R, G, B, A bytes) of one row: luma bytes
// ylo | yhi, pixel k using the chroma of pair k / 2.
template <int OC>
__device__ __forceinline__ void bt601_row8(uint32_t ylo, uint32_t yhi, const Chroma &t,
                                           bool bgr, uint32_t px[kBlockW]) {
#pragma unroll
  for (int k = 0; k < kBlockW; ++k) {
    const uint32_t yv = ((k < 4 ? ylo : yhi) >> (8 * (k & 3))) & 0xFF;
    const int yw = 1220542 * max(static_cast<int>(yv) - 16, 0);
    const int p = k >> 1;
    px[k] = pack_px(sat_u8((yw + t.r[p]) >> 20), sat_u8((yw + t.g[p]) >> 20),
                    sat_u8((yw + t.b[p]) >> 20), bgr, OC == 4);
  }
}

// Stores 8 packed pixels at `out` (the row's pixel x0): 16-byte (RGBA) or
// 8-byte (RGB) non-temporal stores when `vec`, else the first `count`
// pixels byte by byte.
template <int OC>
__device__ __forceinline__ void store_px8(uint8_t *out, const uint32_t px[kBlockW], bool vec,
                                          int count) {
  if (vec) {
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
    for (int k = 0; k < count; ++k)
#pragma unroll
      for (int c = 0; c < OC; ++c) out[k * OC + c] = static_cast<uint8_t>(px[k] >> (8 * c));
  }
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
  const int pairs = min(yp_end, (a.height + 1) / 2);
  for (int yp = yp_begin; yp < pairs; yp += yp_step) {
    const uint8_t *uvrow = a.uv + static_cast<size_t>(yp) * a.uv_stride;
    const bool two = 2 * yp + 1 < a.height;
    // all loads first: the 4 chroma pairs and 8 luma bytes of both rows
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
    const Chroma t = chroma4(uv_lo, uv_hi, a.vu);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int y = 2 * yp + r;
      if (y >= a.height) break;
      uint32_t px[kBlockW];
      bt601_row8<OC>(yw_lo[r], yw_hi[r], t, a.bgr, px);
      uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride + static_cast<size_t>(x0) * OC;
      store_px8<OC>(out, px, full && out_al, min(kBlockW, a.width - x0));
    }
  }
}

// ---------------------------------------------------------------------------
// Planar 4:2:0 (I420 / YV12), interleaved 4:2:2 (YUYV / UYVY / YVYU) and
// interleaved 4:4:4, one `src` buffer each.  Same 8-pixel thread blocks.
//
// 4:2:0 planar follows OpenCV's packed-plane layout, as the reference reads
// it (src/conversions/yuv420p_to_rgb_sc.h:46-110): the two chroma planes
// follow the luma plane, their rows of width/2 bytes two to a source row,
// so chroma row k of a plane starts (k / 2) * stride bytes in, plus the
// plane's first row step when k is odd (width/2, or stride - width/2 for a
// plane that starts half-way along a row).  Pixel x takes chroma element
// x >> 1 of that row.  Reads are clamped to the source extent (rows
// (3h+1)/2), the region the reference's own reads stay in.
__device__ __forceinline__ const uint8_t *chroma_row(const uint8_t *plane, int k, int phase,
                                                     size_t stride, int width) {
  const size_t half = static_cast<size_t>(width / 2);
  const size_t step0 = phase ? stride - half : half;
  return plane + static_cast<size_t>(k / 2) * stride + ((k & 1) ? step0 : 0);
}

// 4 chroma bytes from `p` (index i < n only), as one word.
__device__ __forceinline__ uint32_t chroma_word(const uint8_t *base, const uint8_t *p,
                                                size_t extent, int n) {
  const size_t off = static_cast<size_t>(p - base);
  if (n >= 4 && off + 4 <= extent) {
    uint32_t w;
    __builtin_memcpy(&w, p, 4);
    return w;
  }
  uint32_t w = 0;
  for (int i = 0; i < n && i < 4; ++i)
    if (off + i < extent) w |= static_cast<uint32_t>(p[i]) << (8 * i);
  return w;
}

template <int OC>
__device__ __forceinline__ void yuv420p_rows(const YuvXArgs &a, int bx, int yp_begin,
                                             int yp_step) {
  const int x0 = (bx * kThreads + threadIdx.x) * kBlockW;
  if (x0 >= a.width) return;
  const bool full = x0 + kBlockW <= a.width;
  const bool y_al = ((reinterpret_cast<uintptr_t>(a.src) | a.src_stride) & 7) == 0;
  const bool out_al = ((reinterpret_cast<uintptr_t>(a.dst) | a.dst_stride) &
                       (OC == 4 ? 15 : 7)) == 0;
  const int pairs = (a.height + 1) / 2;
  const int nchroma = (min(a.width - x0, kBlockW) + 1) / 2;
  for (int yp = yp_begin; yp < pairs; yp += yp_step) {
    const bool two = 2 * yp + 1 < a.height;
    const uint8_t *urow = chroma_row(a.u, yp, a.u_phase, a.src_stride, a.width) + x0 / 2;
    const uint8_t *vrow = chroma_row(a.v, yp, a.v_phase, a.src_stride, a.width) + x0 / 2;
    const uint32_t uw = chroma_word(a.src, urow, a.src_extent, nchroma);
    const uint32_t vw = chroma_word(a.src, vrow, a.src_extent, nchroma);
    uint32_t yw_lo[2] = {0, 0}, yw_hi[2] = {0, 0};
    const uint8_t *y0row = a.src + static_cast<size_t>(2 * yp) * a.src_stride + x0;
    if (full && y_al) {
      const uint2 q0 = *reinterpret_cast<const uint2 *>(y0row);
      const uint2 q1 = two ? *reinterpret_cast<const uint2 *>(y0row + a.src_stride) : q0;
      yw_lo[0] = q0.x;
      yw_hi[0] = q0.y;
      yw_lo[1] = q1.x;
      yw_hi[1] = q1.y;
    } else {
      for (int r = 0; r < (two ? 2 : 1); ++r)
        for (int k = 0; k < kBlockW && x0 + k < a.width; ++k) {
          const uint32_t v = y0row[r * a.src_stride + k];
          if (k < 4) yw_lo[r] |= v << (8 * k); else yw_hi[r] |= v << (8 * (k - 4));
        }
    }
    // interleave to U0 V0 U1 V1 | U2 V2 U3 V3
    const uint32_t uv_lo = (uw & 0xFFu) | ((vw & 0xFFu) << 8) | ((uw & 0xFF00u) << 8) |
                           ((vw & 0xFF00u) << 16);
    const uint32_t uv_hi = ((uw >> 16) & 0xFFu) | ((vw >> 8) & 0xFF00u) |
                           ((uw >> 8) & 0xFF0000u) | (vw & 0xFF000000u);
    const Chroma t = chroma4(uv_lo, uv_hi, false);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int y = 2 * yp + r;
      if (y >= a.height) break;
      uint32_t px[kBlockW];
      bt601_row8<OC>(yw_lo[r], yw_hi[r], t, a.bgr, px);
      uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride + static_cast<size_t>(x0) * OC;
      store_px8<OC>(out, px, full && out_al, min(kBlockW, a.width - x0));
    }
  }
}

// 4:2:2: each 4-byte group holds two pixels (Y0 at y_off, Y1 at y_off + 2,
// U at u_off, V at v_off).  src/conversions/yuv422_to_rgb_sc.h: same
// BT.601 limited-range arithmetic as 4:2:0, chroma shared by a pixel pair.
template <int OC>
__device__ __forceinline__ void yuv422_rows(const YuvXArgs &a, int bx, int y_begin,
                                            int y_step) {
  const int x0 = (bx * kThreads + threadIdx.x) * kBlockW;
  if (x0 >= a.width) return;
  const bool full = x0 + kBlockW <= a.width;  // width is even: pairs are whole
  const bool in_al = ((reinterpret_cast<uintptr_t>(a.src) | a.src_stride) & 15) == 0;
  const bool out_al = ((reinterpret_cast<uintptr_t>(a.dst) | a.dst_stride) &
                       (OC == 4 ? 15 : 7)) == 0;
  for (int y = y_begin; y < a.height; y += y_step) {
    const uint8_t *row = a.src + static_cast<size_t>(y) * a.src_stride + 2 * x0;
    uint32_t q[4] = {0, 0, 0, 0};
    if (full && in_al) {
      const uint4 v = *reinterpret_cast<const uint4 *>(row);
      q[0] = v.x;
      q[1] = v.y;
      q[2] = v.z;
      q[3] = v.w;
    } else {
      for (int b = 0; b < 2 * min(kBlockW, a.width - x0); ++b)
        q[b >> 2] |= static_cast<uint32_t>(row[b]) << (8 * (b & 3));
    }
    uint32_t ylo = 0, yhi = 0, uv_lo = 0, uv_hi = 0;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const uint32_t y0 = (q[p] >> (8 * a.y_off)) & 0xFF;
      const uint32_t y1 = (q[p] >> (8 * a.y_off + 16)) & 0xFF;
      const uint32_t u = (q[p] >> (8 * a.u_off)) & 0xFF;
      const uint32_t v = (q[p] >> (8 * a.v_off)) & 0xFF;
      const uint32_t yy = y0 | (y1 << 8), cc = u | (v << 8);
      if (p < 2) {
        ylo |= yy << (16 * p);
        uv_lo |= cc << (16 * p);
      } else {
        yhi |= yy << (16 * (p - 2));
        uv_hi |= cc << (16 * (p - 2));
      }
    }
    const Chroma t = chroma4(uv_lo, uv_hi, false);
    uint32_t px[kBlockW];
    bt601_row8<OC>(ylo, yhi, t, a.bgr, px);
    uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride + static_cast<size_t>(x0) * OC;
    store_px8<OC>(out, px, full && out_al, min(kBlockW, a.width - x0));
  }
}

// 4:4:4 interleaved Y U V, full-range BT.601 in 14-bit fixed point
// (src/conversions/yuv444_coefficients.h; yuv444_to_rgb_sc.h pre-shifts U
// and V by 2 and keeps the top 16 bits of a 32-bit sum, which is the same
// as the >> 14 below):
//   R = Y + ((18678 V' + 2^13) >> 14)
//   G = Y + ((-6472 U' - 9519 V' + 2^13) >> 14)
//   B = Y + ((33292 U' + 2^13) >> 14),  saturated to u8.
template <int OC>
__device__ __forceinline__ void yuv444_rows(const YuvXArgs &a, int bx, int y_begin,
                                            int y_step) {
  const int x0 = (bx * kThreads + threadIdx.x) * kBlockW;
  if (x0 >= a.width) return;
  const bool full = x0 + kBlockW <= a.width;
  const bool in_al = ((reinterpret_cast<uintptr_t>(a.src) | a.src_stride) & 7) == 0;
  const bool out_al = ((reinterpret_cast<uintptr_t>(a.dst) | a.dst_stride) &
                       (OC == 4 ? 15 : 7)) == 0;
  for (int y = y_begin; y < a.height; y += y_step) {
    const uint8_t *row = a.src + static_cast<size_t>(y) * a.src_stride + 3 * x0;
    uint32_t w[6] = {0, 0, 0, 0, 0, 0};
    if (full && in_al) {
      const uint2 *p = reinterpret_cast<const uint2 *>(row);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const uint2 v = p[m];
        w[2 * m] = v.x;
        w[2 * m + 1] = v.y;
      }
    } else {
      for (int b = 0; b < 3 * min(kBlockW, a.width - x0); ++b)
        w[b >> 2] |= static_cast<uint32_t>(row[b]) << (8 * (b & 3));
    }
    auto byte = [&](int i) -> int { return static_cast<int>((w[i >> 2] >> (8 * (i & 3))) & 0xFF); };
    uint32_t px[kBlockW];
#pragma unroll
    for (int k = 0; k < kBlockW; ++k) {
      const int yv = byte(3 * k), u = byte(3 * k + 1) - 128, v = byte(3 * k + 2) - 128;
      px[k] = pack_px(sat_u8(yv + ((18678 * v + (1 << 13)) >> 14)),
                      sat_u8(yv + ((-6472 * u - 9519 * v + (1 << 13)) >> 14)),
                      sat_u8(yv + ((33292 * u + (1 << 13)) >> 14)), a.bgr, OC == 4);
    }
    uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride + static_cast<size_t>(x0) * OC;
    store_px8<OC>(out, px, full && out_al, min(kBlockW, a.width - x0));
  }
}

}  // namespace yuv
}  // namespace kcv

#endif  // KCV_YUV_BODY_H
