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
  const bool al = (reinterpret_cast<uintptr_t>(out) & (OC == 4 ? 15 : 7)) == 0;
  if (vec && !al) {
    // any alignment: plain (not non-temporal) vector stores
    if (OC == 4) {
      reinterpret_cast<u4_any *>(out)[0] = make_uint4(px[0], px[1], px[2], px[3]);
      reinterpret_cast<u4_any *>(out)[1] = make_uint4(px[4], px[5], px[6], px[7]);
      return;
    }
  }
  if (vec && (al || OC == 3)) {
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
      if (al) {
        uint2 *o = reinterpret_cast<uint2 *>(out);
        store_nt(o, make_uint2(w[0], w[1]));
        store_nt(o + 1, make_uint2(w[2], w[3]));
        store_nt(o + 2, make_uint2(w[4], w[5]));
      } else {
        u2_any *o = reinterpret_cast<u2_any *>(out);
        o[0] = make_uint2(w[0], w[1]);
        o[1] = make_uint2(w[2], w[3]);
        o[2] = make_uint2(w[4], w[5]);
      }
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
  const int pairs = min(yp_end, (a.height + 1) / 2);
  for (int yp = yp_begin; yp < pairs; yp += yp_step) {
    const uint8_t *uvrow = a.uv + static_cast<size_t>(yp) * a.uv_stride;
    const bool two = 2 * yp + 1 < a.height;
    // all loads first: the 4 chroma pairs and 8 luma bytes of both rows
    uint32_t uv_lo = 0, uv_hi = 0, yw_lo[2] = {0, 0}, yw_hi[2] = {0, 0};
    if (full) {
      const uint8_t *y0row = a.y + static_cast<size_t>(2 * yp) * a.y_stride + x0;
      const uint2 q = *reinterpret_cast<const u2_any *>(uvrow + x0);
      const uint2 q0 = *reinterpret_cast<const u2_any *>(y0row);
      const uint2 q1 = two ? *reinterpret_cast<const u2_any *>(y0row + a.y_stride) : q0;
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
      store_px8<OC>(out, px, full, min(kBlockW, a.width - x0));
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
    if (full) {
      const uint2 q0 = *reinterpret_cast<const u2_any *>(y0row);
      const uint2 q1 = two ? *reinterpret_cast<const u2_any *>(y0row + a.src_stride) : q0;
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
      store_px8<OC>(out, px, full, min(kBlockW, a.width - x0));
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
  for (int y = y_begin; y < a.height; y += y_step) {
    const uint8_t *row = a.src + static_cast<size_t>(y) * a.src_stride + 2 * x0;
    uint32_t q[4] = {0, 0, 0, 0};
    if (full) {
      const uint4 v = *reinterpret_cast<const u4_any *>(row);
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
    store_px8<OC>(out, px, full, min(kBlockW, a.width - x0));
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
  for (int y = y_begin; y < a.height; y += y_step) {
    const uint8_t *row = a.src + static_cast<size_t>(y) * a.src_stride + 3 * x0;
    uint32_t w[6] = {0, 0, 0, 0, 0, 0};
    if (full) {
      const u2_any *p = reinterpret_cast<const u2_any *>(row);
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
    store_px8<OC>(out, px, full, min(kBlockW, a.width - x0));
  }
}

// ---------------------------------------------------------------------------
// RGB/BGR(A) to YUV, 8 source pixels per thread.
//  4:2:0: 20-bit limited-range weights (yuv42x_coefficients.h:17-29), chroma
//         from the top-left pixel of each 2x2 block (rgb_to_yuv420_sc.h).
//  4:2:2: 14-bit weights, chroma from the pair sums (rgb_to_yuv422_sc.h).
//  4:4:4: 14-bit full-range weights, chroma from B - Y and R - Y of the
//         rounded Y (rgb_to_yuv444_coefficients.h, rgb_to_yuv444_sc.h).

// Loads 8 pixels of SCN bytes at p (the first `count` when not `vec`).
template <int SCN>
__device__ __forceinline__ void load_px8(const uint8_t *p, bool vec, int count,
                                         uint32_t w[2 * SCN]) {
#pragma unroll
  for (int i = 0; i < 2 * SCN; ++i) w[i] = 0;
  if (vec) {
    const u2_any *q = reinterpret_cast<const u2_any *>(p);
#pragma unroll
    for (int i = 0; i < SCN; ++i) {
      const uint2 v = q[i];
      w[2 * i] = v.x;
      w[2 * i + 1] = v.y;
    }
  } else {
    for (int b = 0; b < SCN * count; ++b) w[b >> 2] |= static_cast<uint32_t>(p[b]) << (8 * (b & 3));
  }
}

struct Rgb8 {
  int r[kBlockW], g[kBlockW], b[kBlockW];
};

template <int SCN>
__device__ __forceinline__ Rgb8 split_px8(const uint32_t w[2 * SCN], bool bgr) {
  Rgb8 p;
  auto byte = [&](int i) -> int { return static_cast<int>((w[i >> 2] >> (8 * (i & 3))) & 0xFF); };
#pragma unroll
  for (int k = 0; k < kBlockW; ++k) {
    const int c0 = byte(SCN * k), c1 = byte(SCN * k + 1), c2 = byte(SCN * k + 2);
    p.r[k] = bgr ? c2 : c0;
    p.g[k] = c1;
    p.b[k] = bgr ? c0 : c2;
  }
  return p;
}

// Stores `n` bytes of the words w (byte i of the run is byte i & 3 of w[i >> 2]).
__device__ __forceinline__ void store_bytes(uint8_t *out, const uint32_t *w, int n, bool vec8) {
  if (vec8) {
    if ((reinterpret_cast<uintptr_t>(out) & 7) == 0)
      store_nt(reinterpret_cast<uint2 *>(out), make_uint2(w[0], w[1]));
    else
      *reinterpret_cast<u2_any *>(out) = make_uint2(w[0], w[1]);
    return;
  }
  for (int i = 0; i < n; ++i) out[i] = static_cast<uint8_t>(w[i >> 2] >> (8 * (i & 3)));
}

template <int SCN>
__device__ __forceinline__ void rgb_to_yuv420_rows(const RgbYuvArgs &a, int bx, int yp_begin,
                                                   int yp_step) {
  const int x0 = (bx * kThreads + threadIdx.x) * kBlockW;
  if (x0 >= a.width) return;
  const int count = min(kBlockW, a.width - x0);
  const bool full = count == kBlockW;
  const int pairs = (a.height + 1) / 2;
  const int nchroma = (count + 1) / 2;
  const size_t cw = static_cast<size_t>((a.width + 1) / 2);
  for (int yp = yp_begin; yp < pairs; yp += yp_step) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int y = 2 * yp + r;
      if (y >= a.height) break;
      uint32_t w[2 * SCN];
      load_px8<SCN>(a.src + static_cast<size_t>(y) * a.src_stride + static_cast<size_t>(x0) * SCN,
                    full, count, w);
      const Rgb8 p = split_px8<SCN>(w, a.bgr);
      uint32_t yw[2] = {0, 0};
#pragma unroll
      for (int k = 0; k < kBlockW; ++k) {
        const uint32_t Y = static_cast<uint32_t>(269484 * p.r[k] + 528482 * p.g[k] +
                                                 102760 * p.b[k] + (1 << 19) + (16 << 20)) >> 20;
        yw[k >> 2] |= (Y & 0xFF) << (8 * (k & 3));
      }
      store_bytes(a.dst + static_cast<size_t>(y) * a.dst_stride + x0, yw, count, full);
      if (r == 1) continue;
      // chroma of the block's top-left pixels 0, 2, 4, 6
      uint32_t uw = 0, vw = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int k = 2 * c;
        const int U = (-155188 * p.r[k] - 305135 * p.g[k] + 460324 * p.b[k] + (1 << 19) +
                       (128 << 20)) >> 20;
        const int V = (460324 * p.r[k] - 385875 * p.g[k] - 74448 * p.b[k] + (1 << 19) +
                       (128 << 20)) >> 20;
        uw |= (static_cast<uint32_t>(U) & 0xFF) << (8 * c);
        vw |= (static_cast<uint32_t>(V) & 0xFF) << (8 * c);
      }
      if (a.vu) {
        const uint32_t t = uw;
        uw = vw;
        vw = t;
      }
      if (a.layout == kRgb420SP) {
        // U0 V0 U1 V1 | U2 V2 U3 V3
        const uint32_t lo = (uw & 0xFFu) | ((vw & 0xFFu) << 8) | ((uw & 0xFF00u) << 8) |
                            ((vw & 0xFF00u) << 16);
        const uint32_t hi = ((uw >> 16) & 0xFFu) | ((vw >> 8) & 0xFF00u) |
                            ((uw >> 8) & 0xFF0000u) | (vw & 0xFF000000u);
        const uint32_t uv[2] = {lo, hi};
        uint8_t *out = a.uv + static_cast<size_t>(yp) * a.uv_stride + x0;
        store_bytes(out, uv, 2 * nchroma, nchroma == 4);
      } else {
        // planar: U row h/4 of the chroma area (+ cw when h/2 is odd), V row
        // (h + height + 1)/4 (+ cw when (h + height + 1)/2 is odd), h = 2 yp
        const size_t h0 = static_cast<size_t>(2 * yp);
        const size_t hv = h0 + a.height + 1;
        // Two half rows of cw bytes share a row; for odd widths the second
        // one's last byte lies past column width - 1 (in the row padding, or
        // on the next row's first byte, which that row's own chroma then
        // holds).  Bytes at column >= width are not written.
        const int uhalf = static_cast<int>((h0 / 2) % 2), vhalf = static_cast<int>((hv / 2) % 2);
        uint8_t *urow = a.uv + a.dst_stride * (h0 / 4) + uhalf * cw;
        uint8_t *vrow = a.uv + a.dst_stride * (hv / 4) + vhalf * cw;
        for (int c = 0; c < nchroma; ++c) {
          const size_t j = static_cast<size_t>(x0 / 2 + c);
          if (uhalf * cw + j < static_cast<size_t>(a.width)) urow[j] = static_cast<uint8_t>(uw >> (8 * c));
          if (vhalf * cw + j < static_cast<size_t>(a.width)) vrow[j] = static_cast<uint8_t>(vw >> (8 * c));
        }
      }
    }
  }
}

template <int SCN>
__device__ __forceinline__ void rgb_to_yuv422_rows(const RgbYuvArgs &a, int bx, int y_begin,
                                                   int y_step) {
  const int x0 = (bx * kThreads + threadIdx.x) * kBlockW;
  if (x0 >= a.width) return;
  const int count = min(kBlockW, a.width - x0);  // even: width is even
  const bool full = count == kBlockW;
  for (int y = y_begin; y < a.height; y += y_step) {
    uint32_t w[2 * SCN];
    load_px8<SCN>(a.src + static_cast<size_t>(y) * a.src_stride + static_cast<size_t>(x0) * SCN,
                  full, count, w);
    const Rgb8 p = split_px8<SCN>(w, a.bgr);
    uint32_t q[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int k = 2 * c;
      const int yb = (1 << 13) + (16 << 14), cb = (1 << 13) + (256 << 13);
      const uint32_t y0 = static_cast<uint32_t>(yb + 4211 * p.r[k] + 8258 * p.g[k] + 1606 * p.b[k]) >> 14;
      const uint32_t y1 = static_cast<uint32_t>(yb + 4211 * p.r[k + 1] + 8258 * p.g[k + 1] +
                                                1606 * p.b[k + 1]) >> 14;
      const int r = p.r[k] + p.r[k + 1], g = p.g[k] + p.g[k + 1], b = p.b[k] + p.b[k + 1];
      const uint32_t u = static_cast<uint32_t>(cb - 1212 * r - 2384 * g + 3596 * b) >> 14;
      const uint32_t v = static_cast<uint32_t>(cb + 3596 * r - 3015 * g - 582 * b) >> 14;
      q[c] = ((y0 & 0xFF) << (8 * a.y_off)) | ((y1 & 0xFF) << (8 * a.y_off + 16)) |
             ((u & 0xFF) << (8 * a.u_off)) | ((v & 0xFF) << (8 * a.v_off));
    }
    uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride + 2 * x0;
    if (full && (reinterpret_cast<uintptr_t>(out) & 15) == 0)
      store_nt(reinterpret_cast<uint4 *>(out), make_uint4(q[0], q[1], q[2], q[3]));
    else if (full)
      *reinterpret_cast<u4_any *>(out) = make_uint4(q[0], q[1], q[2], q[3]);
    else
      store_bytes(out, q, 2 * count, false);
  }
}

template <int SCN>
__device__ __forceinline__ void rgb_to_yuv444_rows(const RgbYuvArgs &a, int bx, int y_begin,
                                                   int y_step) {
  const int x0 = (bx * kThreads + threadIdx.x) * kBlockW;
  if (x0 >= a.width) return;
  const int count = min(kBlockW, a.width - x0);
  const bool full = count == kBlockW;
  for (int y = y_begin; y < a.height; y += y_step) {
    uint32_t w[2 * SCN];
    load_px8<SCN>(a.src + static_cast<size_t>(y) * a.src_stride + static_cast<size_t>(x0) * SCN,
                  full, count, w);
    const Rgb8 p = split_px8<SCN>(w, a.bgr);
    uint32_t o[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < kBlockW; ++k) {
      const int Y = (4899 * p.r[k] + 9617 * p.g[k] + 1868 * p.b[k] + (1 << 13)) >> 14;
      const int U = ((128 << 14) + 8061 * (p.b[k] - Y) + (1 << 13)) >> 14;
      const int V = ((128 << 14) + 14369 * (p.r[k] - Y) + (1 << 13)) >> 14;
      const uint32_t c[3] = {sat_u8(Y), sat_u8(U), sat_u8(V)};
#pragma unroll
      for (int j = 0; j < 3; ++j) o[(3 * k + j) >> 2] |= c[j] << (8 * ((3 * k + j) & 3));
    }
    uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride + 3 * x0;
    if (full && (reinterpret_cast<uintptr_t>(out) & 7) == 0) {
      uint2 *q = reinterpret_cast<uint2 *>(out);
      store_nt(q, make_uint2(o[0], o[1]));
      store_nt(q + 1, make_uint2(o[2], o[3]));
      store_nt(q + 2, make_uint2(o[4], o[5]));
    } else if (full) {
      u2_any *q = reinterpret_cast<u2_any *>(out);
      q[0] = make_uint2(o[0], o[1]);
      q[1] = make_uint2(o[2], o[3]);
      q[2] = make_uint2(o[4], o[5]);
    } else {
      store_bytes(out, o, 3 * count, false);
    }
  }
}

}  // namespace yuv
}  // namespace kcv

#endif  // KCV_YUV_BODY_H
