// Device bodies of the geometric gathers for gfx950 (per-image kernels in
// transform.hip, batched kernels in batch.hip): remap with s16 + 5/5-bit fractional maps,
// remap with float maps and warp-perspective.
//
// Work layout: each wavefront owns 8 x 8 destination pixels per step, so the
// source footprint of one load instruction stays a compact patch (a few
// cache lines) under rotation or displacement instead of a long slanted run.
// A workgroup of 4 wavefronts covers 32 x 32 pixels; every thread handles
// four pixels from four different 8 x 8 tiles and issues all of their loads
// before using any result.  Four-channel pixels are fetched as one dword per
// corner.
//
// Float semantics follow the reference's AArch64 build: the lerps and the
// row/column homography updates are fused multiply-adds (vmlaq_f32 /
// vmlaq_n_f32 and scalar a*b+c under clang's default contraction), 1/w is a
// correctly rounded division, conversions are the saturating FCVTM*/FCVTA*
// forms (NaN -> 0).  This file is compiled with -ffp-contract=off so only
// the explicit fmaf calls fuse.
//
// Reference: src/transform/remap_s16point5_neon.cpp:16-1262,
// src/transform/remap_f32_neon.cpp:14-220, src/transform/transform_neon.h,
// src/transform/warp_perspective_neon.cpp:35-193.

#ifndef KCV_TRANSFORM_BODY_H
#define KCV_TRANSFORM_BODY_H

#include "kcv_internal.h"

namespace kcv {
namespace xform {

constexpr int kThreads = 256;
constexpr int kBlockPx = 32;  // workgroup tile edge in pixels
constexpr int kPerThread = 4;

// Pixel i (0..3) of this thread: tile (wave*4 + i) of the 4 x 4 grid of
// 8 x 8 tiles, lane -> position inside the tile.
__device__ __forceinline__ void pixel_of(int i, int bx, int by, int &x, int &y) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = wave * kPerThread + i;
  x = bx * kBlockPx + (j & 3) * 8 + (lane & 7);
  y = by * kBlockPx + (j >> 2) * 8 + (lane >> 3);
}

// Up to four channel bytes of one pixel packed in a dword.
template <int C>
__device__ __forceinline__ uint32_t load_px(const uint8_t *p, bool aligned4) {
  if (C == 4 && aligned4) return *reinterpret_cast<const uint32_t *>(p);
  uint32_t v = 0;
#pragma unroll
  for (int c = 0; c < C; ++c) v |= static_cast<uint32_t>(p[c]) << (8 * c);
  return v;
}

template <int C>
__device__ __forceinline__ void store_px(uint8_t *p, uint32_t v, bool aligned4) {
  if (C == 4 && aligned4) {
    *reinterpret_cast<uint32_t *>(p) = v;
    return;
  }
#pragma unroll
  for (int c = 0; c < C; ++c) p[c] = static_cast<uint8_t>(v >> (8 * c));
}

__device__ __forceinline__ uint32_t byte_of(uint32_t v, int c) {
  return (v >> (8 * c)) & 0xFFu;
}

struct __attribute__((aligned(4))) U32x2 {
  uint32_t x, y;
};

// The two horizontal taps of one source row, pixel x0 and (when `adjacent`)
// pixel x0 + 1, else x0 again, with one load where possible: an 8-byte load
// of both 4-channel pixels, or for 1 channel the 8 bytes from the dword
// holding x0.  `tail_ok` says those 8 bytes lie inside the image: within
// this row, or (rows below this one exist and the width is at least 8)
// running into the next row.
template <int C>
__device__ __forceinline__ void load_pair(const uint8_t *row, uint32_t x0, bool adjacent,
                                          bool tail_ok, bool aligned4, uint32_t &pa,
                                          uint32_t &pb) {
  if (C == 4 && aligned4 && adjacent) {
    const U32x2 v = *reinterpret_cast<const U32x2 *>(row + 4 * static_cast<size_t>(x0));
    pa = v.x;
    pb = v.y;
    return;
  }
  if (C == 1 && aligned4 && tail_ok) {
    const U32x2 v = *reinterpret_cast<const U32x2 *>(row + (x0 & ~3u));
    const uint64_t q = (static_cast<uint64_t>(v.y) << 32) | v.x;
    const uint32_t sh = 8 * (x0 & 3u);
    pa = static_cast<uint32_t>(q >> sh) & 0xFFu;
    pb = adjacent ? static_cast<uint32_t>(q >> (sh + 8)) & 0xFFu : pa;
    return;
  }
  pa = load_px<C>(row + static_cast<size_t>(x0) * C, aligned4);
  pb = adjacent ? load_px<C>(row + static_cast<size_t>(x0 + 1) * C, aligned4) : pa;
}

// ---------------------------------------------------------------------------
// remap s16point5
__device__ __forceinline__ uint32_t bilinear5(uint32_t a, uint32_t b,
                                              uint32_t c, uint32_t d,
                                              uint32_t fx, uint32_t fy) {
  const uint32_t line0 = (32 - fx) * a + fx * b;
  const uint32_t line1 = (32 - fx) * c + fx * d;
  return (line0 * (32 - fy) + line1 * fy + 512) >> 10;
}

// Destination tile (bx, by) of kBlockPx x kBlockPx pixels.
template <int C, int BORDER>
__device__ __forceinline__ void remap_s16point5_tile(const RemapS16Args &a, int bx, int by) {
  const bool src_al = ((reinterpret_cast<uintptr_t>(a.src) | a.src_stride) & 3) == 0;
  const bool dst_al = ((reinterpret_cast<uintptr_t>(a.dst) | a.dst_stride) & 3) == 0;
  const bool map_al = ((reinterpret_cast<uintptr_t>(a.mapxy) | a.mapxy_stride) & 3) == 0;
  const int xmax = a.src_width - 1, ymax = a.src_height - 1;
  uint32_t bv = 0;
#pragma unroll
  for (int c = 0; c < C; ++c)
    bv |= static_cast<uint32_t>(a.border_value[C == 4 ? c : 0]) << (8 * c);
  {
    int px[kPerThread], py[kPerThread];
    bool ok[kPerThread];
    uint32_t xy[kPerThread], fr[kPerThread];
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      pixel_of(i, bx, by, px[i], py[i]);
      ok[i] = px[i] < a.dst_width && py[i] < a.dst_height;
      const int x = ok[i] ? px[i] : 0, y = ok[i] ? py[i] : 0;
      const uint16_t *m = reinterpret_cast<const uint16_t *>(
          reinterpret_cast<const uint8_t *>(a.mapxy) + y * a.mapxy_stride + 4 * x);
      xy[i] = map_al ? *reinterpret_cast<const uint32_t *>(m)
                     : (static_cast<uint32_t>(m[0]) | (static_cast<uint32_t>(m[1]) << 16));
      fr[i] = *reinterpret_cast<const uint16_t *>(
          reinterpret_cast<const uint8_t *>(a.mapfrac) + y * a.mapfrac_stride + 2 * x);
    }
    uint32_t pa[kPerThread], pb[kPerThread], pc[kPerThread], pd[kPerThread];
    uint32_t fx[kPerThread], fy[kPerThread];
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      const int mx = static_cast<int16_t>(xy[i] & 0xFFFF);
      const int my = static_cast<int16_t>(xy[i] >> 16);
      fx[i] = fr[i] & 31u;
      fy[i] = (fr[i] >> 5) & 31u;
      if (BORDER == kReplicate) {
        // RemapS16Point5Replicate (:16-137) / get_coordinates_replicate
        // (:621-640): fractions of negative coordinates are dropped, taps
        // clamp to the image.
        if (mx < 0) fx[i] = 0;
        if (my < 0) fy[i] = 0;
        const int x0 = min(max(mx, 0), xmax), y0 = min(max(my, 0), ymax);
        const int y1 = y0 + (y0 < ymax ? 1 : 0);
        const bool adj = x0 < xmax;
        const bool tail = (x0 & ~3) + 7 <= xmax;
        const uint8_t *r0 = a.src + static_cast<size_t>(y0) * a.src_stride;
        const uint8_t *r1 = a.src + static_cast<size_t>(y1) * a.src_stride;
        load_pair<C>(r0, x0, adj, tail || (y0 < ymax && xmax >= 7), src_al, pa[i], pb[i]);
        load_pair<C>(r1, x0, adj, tail || (y1 < ymax && xmax >= 7), src_al, pc[i], pd[i]);
      } else {
        // RemapS16Point5ConstantBorder (:321-425) / get_coordinates_constant
        // (:860-881): each corner tests its own range with 16-bit unsigned
        // compares and takes the border value when outside.
        const uint32_t ux0 = static_cast<uint16_t>(mx), uy0 = static_cast<uint16_t>(my);
        const uint32_t ux1 = static_cast<uint16_t>(ux0 + 1);
        const uint32_t uy1 = static_cast<uint16_t>(uy0 + 1);
        const bool x0in = ux0 < static_cast<uint32_t>(a.src_width);
        const bool x1in = ux1 < static_cast<uint32_t>(a.src_width);
        const bool y0in = uy0 < static_cast<uint32_t>(a.src_height);
        const bool y1in = uy1 < static_cast<uint32_t>(a.src_height);
        const uint32_t ry0 = y0in ? uy0 : 0, ry1 = y1in ? uy1 : 0;
        const uint8_t *r0 = a.src + static_cast<size_t>(ry0) * a.src_stride;
        const uint8_t *r1 = a.src + static_cast<size_t>(ry1) * a.src_stride;
        if (x0in && x1in) {
          // both columns inside: one paired load per row that is inside
          const bool tail = (ux0 & ~3u) + 7 <= static_cast<uint32_t>(xmax);
          uint32_t t0, t1, b0, b1;
          load_pair<C>(r0, ux0, true, tail || (ry0 < static_cast<uint32_t>(ymax) && xmax >= 7),
                      src_al, t0, t1);
          load_pair<C>(r1, ux0, true, tail || (ry1 < static_cast<uint32_t>(ymax) && xmax >= 7),
                      src_al, b0, b1);
          pa[i] = y0in ? t0 : bv;
          pb[i] = y0in ? t1 : bv;
          pc[i] = y1in ? b0 : bv;
          pd[i] = y1in ? b1 : bv;
        } else {
          pa[i] = (x0in && y0in) ? load_px<C>(r0 + ux0 * C, src_al) : bv;
          pb[i] = (x1in && y0in) ? load_px<C>(r0 + ux1 * C, src_al) : bv;
          pc[i] = (x0in && y1in) ? load_px<C>(r1 + ux0 * C, src_al) : bv;
          pd[i] = (x1in && y1in) ? load_px<C>(r1 + ux1 * C, src_al) : bv;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      if (!ok[i]) continue;
      uint32_t out = 0;
#pragma unroll
      for (int c = 0; c < C; ++c)
        out |= min(bilinear5(byte_of(pa[i], c), byte_of(pb[i], c), byte_of(pc[i], c),
                             byte_of(pd[i], c), fx[i], fy[i]), 255u) << (8 * c);
      store_px<C>(a.dst + static_cast<size_t>(py[i]) * a.dst_stride + px[i] * C, out, dst_al);
    }
  }
}

// ---------------------------------------------------------------------------
// Float gathers (transform_neon.h:189-491).
__device__ __forceinline__ uint32_t cvtm_u32(float f) {
  if (!(f == f)) return 0;
  const float r = floorf(f);
  if (r <= 0.0f) return 0;
  if (r >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(r);
}

__device__ __forceinline__ int32_t cvtm_s32(float f) {
  if (!(f == f)) return 0;
  const float r = floorf(f);
  if (r >= 2147483648.0f) return 0x7FFFFFFF;
  if (r < -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(r);
}

__device__ __forceinline__ uint32_t cvta_u32(float f) {
  if (!(f == f)) return 0;
  const float r = roundf(f);  // ties away from zero
  if (r <= 0.0f) return 0;
  if (r >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(r);
}

__device__ __forceinline__ int32_t cvta_s32(float f) {
  if (!(f == f)) return 0;
  const float r = roundf(f);
  if (r >= 2147483648.0f) return 0x7FFFFFFF;
  if (r < -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(r);
}

struct Src {
  const uint8_t *p;
  size_t stride;
  uint32_t xmax, ymax;
  bool aligned4;
  uint32_t bv;  // border value bytes, one per channel
};

// Gathered corners of one pixel, ready for the lerp.
struct Quad {
  uint32_t a, b, c, d;
  float xfrac, yfrac;
};

// load_quad_pixels_{replicate,constant} (transform_neon.h:189-329).
template <int C, int BORDER>
__device__ __forceinline__ Quad gather_linear(const Src &s, float xf, float yf) {
  Quad q;
  q.xfrac = xf - floorf(xf);
  q.yfrac = yf - floorf(yf);
  if (BORDER == kReplicate) {
    const uint32_t x0 = min(cvtm_u32(xf), s.xmax);
    const uint32_t y0 = min(cvtm_u32(yf), s.ymax);
    const uint32_t x1 = x0 + ((xf >= 0.0f && x0 < s.xmax) ? 1u : 0u);
    const uint32_t y1 = y0 + ((yf >= 0.0f && y0 < s.ymax) ? 1u : 0u);
    const uint8_t *r0 = s.p + static_cast<size_t>(y0) * s.stride;
    const uint8_t *r1 = s.p + static_cast<size_t>(y1) * s.stride;
    const bool adj = x1 != x0;
    const bool tail = (x0 & ~3u) + 7 <= s.xmax;
    load_pair<C>(r0, x0, adj, tail || (y0 < s.ymax && s.xmax >= 7), s.aligned4, q.a, q.b);
    load_pair<C>(r1, x0, adj, tail || (y1 < s.ymax && s.xmax >= 7), s.aligned4, q.c, q.d);
  } else {
    const uint32_t x0 = static_cast<uint32_t>(cvtm_s32(xf));
    const uint32_t y0 = static_cast<uint32_t>(cvtm_s32(yf));
    const uint32_t x1 = x0 + 1u, y1 = y0 + 1u;
    const bool x0in = x0 <= s.xmax, x1in = x1 <= s.xmax;
    const bool y0in = y0 <= s.ymax, y1in = y1 <= s.ymax;
    const uint32_t ry0 = y0in ? y0 : 0, ry1 = y1in ? y1 : 0;
    const uint8_t *r0 = s.p + static_cast<size_t>(ry0) * s.stride;
    const uint8_t *r1 = s.p + static_cast<size_t>(ry1) * s.stride;
    if (x0in && x1in) {
      const bool tail = (x0 & ~3u) + 7 <= s.xmax;
      uint32_t t0, t1, b0, b1;
      load_pair<C>(r0, x0, true, tail || (ry0 < s.ymax && s.xmax >= 7), s.aligned4, t0, t1);
      load_pair<C>(r1, x0, true, tail || (ry1 < s.ymax && s.xmax >= 7), s.aligned4, b0, b1);
      q.a = y0in ? t0 : s.bv;
      q.b = y0in ? t1 : s.bv;
      q.c = y1in ? b0 : s.bv;
      q.d = y1in ? b1 : s.bv;
    } else {
      q.a = (x0in && y0in) ? load_px<C>(r0 + static_cast<size_t>(x0) * C, s.aligned4) : s.bv;
      q.b = (x1in && y0in) ? load_px<C>(r0 + static_cast<size_t>(x1) * C, s.aligned4) : s.bv;
      q.c = (x0in && y1in) ? load_px<C>(r1 + static_cast<size_t>(x0) * C, s.aligned4) : s.bv;
      q.d = (x1in && y1in) ? load_px<C>(r1 + static_cast<size_t>(x1) * C, s.aligned4) : s.bv;
    }
  }
  return q;
}

// FCVTAU of a lerp result: round half away from zero, NaN to 0.  A lerp of
// bytes with a fraction in [0, 1) is NaN or lies in [0, 255] (the fused
// steps cannot round outside the range of their inputs), so no saturation
// test is needed; the NaN test also covers signalling NaNs.
__device__ __forceinline__ uint32_t round_lerp_u8(float v) {
  const float c = v == v ? v : 0.0f;
  const float t = truncf(c);
  return static_cast<uint32_t>(t) + (c - t >= 0.5f ? 1u : 0u);
}

// lerp_2d (transform_neon.h:331-337) per channel.
template <int C>
__device__ __forceinline__ uint32_t lerp_quad(const Quad &q) {
  uint32_t out = 0;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const float pa = static_cast<float>(byte_of(q.a, c));
    const float pb = static_cast<float>(byte_of(q.b, c));
    const float pc = static_cast<float>(byte_of(q.c, c));
    const float pd = static_cast<float>(byte_of(q.d, c));
    const float line0 = fmaf(pb - pa, q.xfrac, pa);
    const float line1 = fmaf(pd - pc, q.xfrac, pc);
    const float res = fmaf(line1 - line0, q.yfrac, line0);
    out |= round_lerp_u8(res) << (8 * c);
  }
  return out;
}

// transform_pixels_{replicate,constant} (transform_neon.h:339-491).
template <int C, int BORDER>
__device__ __forceinline__ uint32_t gather_nearest(const Src &s, float xf, float yf) {
  if (BORDER == kReplicate) {
    const uint32_t x = min(cvta_u32(xf), s.xmax);
    const uint32_t y = min(cvta_u32(yf), s.ymax);
    return load_px<C>(s.p + static_cast<size_t>(y) * s.stride + static_cast<size_t>(x) * C,
                      s.aligned4);
  }
  const uint32_t x = static_cast<uint32_t>(cvta_s32(xf));
  const uint32_t y = static_cast<uint32_t>(cvta_s32(yf));
  if (x <= s.xmax && y <= s.ymax)
    return load_px<C>(s.p + static_cast<size_t>(y) * s.stride + static_cast<size_t>(x) * C,
                      s.aligned4);
  return s.bv;
}

template <int C>
__device__ __forceinline__ Src make_src(const uint8_t *p, size_t stride, int w, int h,
                                        const uint8_t *border_value) {
  Src s;
  s.p = p;
  s.stride = stride;
  s.xmax = static_cast<uint32_t>(w - 1);
  s.ymax = static_cast<uint32_t>(h - 1);
  s.aligned4 = ((reinterpret_cast<uintptr_t>(p) | stride) & 3) == 0;
  s.bv = 0;
#pragma unroll
  for (int c = 0; c < C; ++c) s.bv |= static_cast<uint32_t>(border_value[c]) << (8 * c);
  return s;
}

// Shared body: `coords(x, y, xf, yf)` gives the source position of a pixel.
template <int C, int INTERP, int BORDER, class Coords>
__device__ __forceinline__ void gather_tiles(const Src &s, uint8_t *dst, size_t dst_stride,
                                             int dst_width, int dst_height, int bx,
                                             int by, const Coords &coords) {
  const bool dst_al = ((reinterpret_cast<uintptr_t>(dst) | dst_stride) & 3) == 0;
  int px[kPerThread], py[kPerThread];
  bool ok[kPerThread];
  float xf[kPerThread], yf[kPerThread];
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    pixel_of(i, bx, by, px[i], py[i]);
    ok[i] = px[i] < dst_width && py[i] < dst_height;
    coords(ok[i] ? px[i] : 0, ok[i] ? py[i] : 0, xf[i], yf[i]);
  }
  uint32_t out[kPerThread];
  if (INTERP == 1) {
    Quad q[kPerThread];
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) q[i] = gather_linear<C, BORDER>(s, xf[i], yf[i]);
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) out[i] = lerp_quad<C>(q[i]);
  } else {
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) out[i] = gather_nearest<C, BORDER>(s, xf[i], yf[i]);
  }
#pragma unroll
  for (int i = 0; i < kPerThread; ++i)
    if (ok[i])
      store_px<C>(dst + static_cast<size_t>(py[i]) * dst_stride + px[i] * C, out[i], dst_al);
}

template <int C, int INTERP, int BORDER>
__device__ __forceinline__ void remap_f32_tile(const RemapF32Args &a, int bx, int by) {
  const Src s = make_src<C>(a.src, a.src_stride, a.src_width, a.src_height, a.border_value);
  auto coords = [&](int x, int y, float &xf, float &yf) {
    xf = *reinterpret_cast<const float *>(reinterpret_cast<const uint8_t *>(a.mapx) +
                                          y * a.mapx_stride + 4 * x);
    yf = *reinterpret_cast<const float *>(reinterpret_cast<const uint8_t *>(a.mapy) +
                                          y * a.mapy_stride + 4 * x);
  };
  gather_tiles<C, INTERP, BORDER>(s, a.dst, a.dst_stride, a.dst_width, a.dst_height, bx, by,
                                  coords);
}

template <int C, int INTERP, int BORDER>
__device__ __forceinline__ void warp_tile(const WarpArgs &a, int bx, int by) {
  const Src s = make_src<C>(a.src, a.src_stride, a.src_width, a.src_height, a.border_value);
  // warp_perspective_neon.cpp:55-76 and :98-102
  auto coords = [&](int x, int y, float &xf, float &yf) {
    const float fy = static_cast<float>(y);
    const float tx0 = fmaf(a.t[1], fy, a.t[2]);
    const float ty0 = fmaf(a.t[4], fy, a.t[5]);
    const float tw0 = fmaf(a.t[7], fy, a.t[8]);
    const float fx = static_cast<float>(static_cast<uint32_t>(x));
    const float tx = fmaf(fx, a.t[0], tx0);
    const float ty = fmaf(fx, a.t[3], ty0);
    const float tw = fmaf(fx, a.t[6], tw0);
    const float iw = (a.check_zero && tw == 0.0f) ? 0.0f : 1.0f / tw;
    xf = tx * iw;
    yf = ty * iw;
  };
  gather_tiles<C, INTERP, BORDER>(s, a.dst, a.dst_stride, a.dst_width, a.dst_height, bx, by,
                                  coords);
}

}  // namespace xform
}  // namespace kcv

#endif  // KCV_TRANSFORM_BODY_H
