// Device bodies of the geometric gathers for gfx950 (per-image kernels in
// transform.hip, batched kernels in batch.hip): remap with s16 + 5/5-bit fractional maps,
// remap with float maps and warp-perspective.
//
// Work layout: a workgroup of 256 threads covers a 64 x 16 tile of
// destination pixels and each thread owns 4 horizontally adjacent pixels of
// one row (16 threads per tile row).  Map entries are then read with 16-byte
// loads and results written with one 4-16 byte store per thread, and the
// gathers of a thread are independent loads with no branch between them.
// Each source row of a bilinear tap is one load of a pixel pair.
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

// ---------------------------------------------------------------------------
// remap s16point5
//
// Tiles of 64 x 16 destination pixels; a thread owns 4 horizontally adjacent
// pixels (16 threads per tile row), so the map entries of a thread are one
// 16-byte and one 8-byte load and its results one 4- or 16-byte store.  The
// gathers are branch-free: each source row contributes one load covering the
// pixel pair (cx, cx + 1), with cx clamped so the pair lies inside the row,
// and taps outside the image are then replaced by the border value.  With
// no branches between them all eight gathers of a thread are in flight
// together.  Source offsets fit 31 bits (stride <= 65535 / channels and
// coordinates <= 32767 are preconditions), so the gathers are buffer loads
// with a 32-bit offset.
constexpr uint32_t kRawBufferWord3 = 0x00020000u;  // gfx9 raw buffer resource

// A source image as a raw buffer resource read through aligned windows.
// The resource starts at the image's first byte rounded down to a dword
// (`lead` bytes before it) and covers whole dwords up to the one holding
// the image's last byte (`records` bytes).  Every dword a load fetches
// therefore holds at least one image byte and lies in a mapped page; dwords
// of a window beyond `records` read as 0 without a memory access (buffer
// loads of several dwords are range checked per dword).  Byte-granular
// loads at odd offsets cost the texture units more than an aligned load of
// the same bytes (the 4K one-channel s16point5 remap: 24.5 us with 2-byte
// tap pairs, 22.1 with aligned 8-byte windows), so the tap bytes of a
// gather are cut out of an aligned window instead.
struct SrcWindow {
  __amdgpu_buffer_rsrc_t rsrc;
  uint32_t lead, records;
};

// Largest image span read through a buffer resource: offsets plus the lead
// and the dword rounding stay below 2^31.
constexpr uint64_t kBufferSpanLimit = (1ull << 31) - 8;

// `span` bytes from `p` (span < kBufferSpanLimit).
__device__ __forceinline__ SrcWindow make_window(const uint8_t *p, uint64_t span) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  SrcWindow w;
  w.lead = static_cast<uint32_t>(a & 3u);
  w.records = static_cast<uint32_t>((span + w.lead + 3) & ~uint64_t{3});
  w.rsrc = __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(a - w.lead), 0,
                                             static_cast<int>(w.records), kRawBufferWord3);
  return w;
}

// Bytes off .. off + 3 of the image (those up to its last byte are valid)
// from the aligned 8-byte window at the dword holding byte off.  A window
// may reach past the resource's end: multi-dword buffer loads are range
// checked per dword, so such a dword reads as 0 and is not fetched, and the
// dwords holding wanted bytes are always inside.
__device__ __forceinline__ uint32_t window4(const SrcWindow &w, uint32_t off) {
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
  const uint32_t o = off + w.lead;
  const u32x2 v = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(
                                                w.rsrc, static_cast<int>(o & ~3u), 0, 0));
  return __builtin_amdgcn_alignbyte(v.y, v.x, o & 3u);
}

// Bytes off .. off + 5 (the first four in lo, the next two in hi) from the
// aligned 12-byte window at the dword holding byte off.
__device__ __forceinline__ void window6(const SrcWindow &w, uint32_t off, uint32_t &lo,
                                        uint32_t &hi) {
  typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));
  const uint32_t o = off + w.lead;
  const u32x3 v = __builtin_bit_cast(u32x3, __builtin_amdgcn_raw_buffer_load_b96(
                                                w.rsrc, static_cast<int>(o & ~3u), 0, 0));
  lo = __builtin_amdgcn_alignbyte(v.y, v.x, o & 3u);
  hi = __builtin_amdgcn_alignbyte(v.z, v.y, o & 3u) & 0xFFFFu;
}

__device__ __forceinline__ uint32_t bilinear5(uint32_t a, uint32_t b,
                                              uint32_t c, uint32_t d,
                                              uint32_t fx, uint32_t fy) {
  const uint32_t line0 = (32 - fx) * a + fx * b;
  const uint32_t line1 = (32 - fx) * c + fx * d;
  return (line0 * (32 - fy) + line1 * fy + 512) >> 10;
}

// Taps of one destination pixel from its map entries (xy: x in the low, y
// in the high 16 bits; fr: 5-bit x and y fractions): the pair column cx
// (taps cx + i0 and cx + i1 of rows r0 and r1), the fractions and, for the
// constant border, which taps lie inside the image.  Rows are clamped into
// the image for both borders (a constant-border tap outside it is replaced
// by the border value, whatever row it reads).
struct S16Tap {
  int cx, r0, r1, i0, i1;
  uint32_t fx, fy;
  bool ina, inb, inc, ind;
};

template <int BORDER>
__device__ __forceinline__ S16Tap s16_tap(int w, int h, uint32_t xy, uint32_t fr) {
  const int mx = static_cast<int16_t>(xy & 0xFFFF);
  const int my = static_cast<int16_t>(xy >> 16);
  S16Tap t;
  t.fx = fr & 31u;
  t.fy = (fr >> 5) & 31u;
  t.ina = t.inb = t.inc = t.ind = true;
  t.cx = max(min(mx, w - 2), 0);
  t.r0 = min(max(my, 0), h - 1);
  if (BORDER == kReplicate) {
    // RemapS16Point5Replicate (:16-137) / get_coordinates_replicate
    // (:621-640): fractions of negative coordinates are dropped, taps
    // clamp to the image.
    if (mx < 0) t.fx = 0;
    if (my < 0) t.fy = 0;
    const int x0 = min(max(mx, 0), w - 1);
    t.r1 = min(t.r0 + 1, h - 1);
    t.i0 = x0 - t.cx;
    t.i1 = min(x0 + 1, w - 1) - t.cx;
  } else {
    // RemapS16Point5ConstantBorder (:321-425) / get_coordinates_constant
    // (:860-881): each corner tests its own range with 16-bit unsigned
    // compares and takes the border value when outside.
    const uint32_t ux0 = static_cast<uint16_t>(mx), uy0 = static_cast<uint16_t>(my);
    const uint32_t ux1 = static_cast<uint16_t>(ux0 + 1);
    const uint32_t uy1 = static_cast<uint16_t>(uy0 + 1);
    const bool x0in = ux0 < static_cast<uint32_t>(w), x1in = ux1 < static_cast<uint32_t>(w);
    const bool y0in = uy0 < static_cast<uint32_t>(h), y1in = uy1 < static_cast<uint32_t>(h);
    t.ina = x0in && y0in;
    t.inb = x1in && y0in;
    t.inc = x0in && y1in;
    t.ind = x1in && y1in;
    t.r1 = min(max(my + 1, 0), h - 1);
    t.i0 = (mx - t.cx) & 1;  // meaningful only when the tap is inside
    t.i1 = (mx + 1 - t.cx) & 1;
  }
  return t;
}

// The pixel from its four tap values (C channels each).
template <int C, int BORDER>
__device__ __forceinline__ uint32_t s16_blend(const S16Tap &t, uint32_t pa, uint32_t pb,
                                              uint32_t pc, uint32_t pd, uint32_t bv) {
  if (BORDER != kReplicate) {
    pa = t.ina ? pa : bv;
    pb = t.inb ? pb : bv;
    pc = t.inc ? pc : bv;
    pd = t.ind ? pd : bv;
  }
  uint32_t out = 0;
#pragma unroll
  for (int c = 0; c < C; ++c)
    out |= bilinear5(byte_of(pa, c), byte_of(pb, c), byte_of(pc, c), byte_of(pd, c), t.fx, t.fy)
           << (8 * c);
  return out;
}

// One destination pixel, its taps gathered from memory.
template <int C, int BORDER>
__device__ __forceinline__ uint32_t remap_s16_px(const SrcWindow &src, uint32_t stride,
                                                 int w, int h, uint32_t xy, uint32_t fr,
                                                 uint32_t bv) {
  const S16Tap t = s16_tap<BORDER>(w, h, xy, fr);
  const uint32_t off0 = static_cast<uint32_t>(t.r0) * stride + static_cast<uint32_t>(t.cx * C);
  const uint32_t off1 = static_cast<uint32_t>(t.r1) * stride + static_cast<uint32_t>(t.cx * C);
  uint32_t pa, pb, pc, pd;
  if (C == 1) {
    const uint32_t t0 = window4(src, off0);
    const uint32_t t1 = window4(src, off1);
    pa = (t0 >> (8 * t.i0)) & 0xFFu;
    pb = (t0 >> (8 * t.i1)) & 0xFFu;
    pc = (t1 >> (8 * t.i0)) & 0xFFu;
    pd = (t1 >> (8 * t.i1)) & 0xFFu;
  } else {
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 t0 = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(
                                                    src.rsrc, off0 + src.lead, 0, 0));
    const u32x2 t1 = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(
                                                    src.rsrc, off1 + src.lead, 0, 0));
    pa = t.i0 ? t0.y : t0.x;
    pb = t.i1 ? t0.y : t0.x;
    pc = t.i0 ? t1.y : t1.x;
    pd = t.i1 ? t1.y : t1.x;
  }
  return s16_blend<C, BORDER>(t, pa, pb, pc, pd, bv);
}

// Same pixel for a source one pixel wide (no pair fits in a row).
template <int C, int BORDER>
__device__ __forceinline__ uint32_t remap_s16_px_narrow(const RemapS16Args &a, uint32_t xy,
                                                        uint32_t fr, uint32_t bv) {
  const int mx = static_cast<int16_t>(xy & 0xFFFF);
  const int my = static_cast<int16_t>(xy >> 16);
  uint32_t fx = fr & 31u, fy = (fr >> 5) & 31u;
  const int h = a.src_height;
  uint32_t pa, pb, pc, pd;
  if (BORDER == kReplicate) {
    if (mx < 0) fx = 0;
    if (my < 0) fy = 0;
    const int y0 = min(max(my, 0), h - 1), y1 = y0 + (y0 < h - 1 ? 1 : 0);
    pa = pb = load_px<C>(a.src + static_cast<size_t>(y0) * a.src_stride, false);
    pc = pd = load_px<C>(a.src + static_cast<size_t>(y1) * a.src_stride, false);
  } else {
    const uint32_t ux0 = static_cast<uint16_t>(mx), uy0 = static_cast<uint16_t>(my);
    const uint32_t ux1 = static_cast<uint16_t>(ux0 + 1), uy1 = static_cast<uint16_t>(uy0 + 1);
    const uint8_t *r0 = a.src + static_cast<size_t>(uy0 < static_cast<uint32_t>(h) ? uy0 : 0) * a.src_stride;
    const uint8_t *r1 = a.src + static_cast<size_t>(uy1 < static_cast<uint32_t>(h) ? uy1 : 0) * a.src_stride;
    const bool y0in = uy0 < static_cast<uint32_t>(h), y1in = uy1 < static_cast<uint32_t>(h);
    pa = (ux0 == 0 && y0in) ? load_px<C>(r0, false) : bv;
    pb = (ux1 == 0 && y0in) ? load_px<C>(r0, false) : bv;
    pc = (ux0 == 0 && y1in) ? load_px<C>(r1, false) : bv;
    pd = (ux1 == 0 && y1in) ? load_px<C>(r1, false) : bv;
  }
  uint32_t out = 0;
#pragma unroll
  for (int c = 0; c < C; ++c)
    out |= bilinear5(byte_of(pa, c), byte_of(pb, c), byte_of(pc, c), byte_of(pd, c), fx, fy)
           << (8 * c);
  return out;
}

// Map entries of the pixels x .. x + 3 of row y (n of them inside the
// destination; the others repeat pixel x).
__device__ __forceinline__ void load_s16_maps(const RemapS16Args &a, int x, int y, int n,
                                              uint32_t (&xy)[4], uint32_t (&fr)[4]) {
  const uint8_t *mrow = reinterpret_cast<const uint8_t *>(a.mapxy) + static_cast<size_t>(y) * a.mapxy_stride;
  const uint8_t *frow = reinterpret_cast<const uint8_t *>(a.mapfrac) + static_cast<size_t>(y) * a.mapfrac_stride;
  if (n == 4 && ((reinterpret_cast<uintptr_t>(mrow) & 15) | (reinterpret_cast<uintptr_t>(frow) & 7)) == 0) {
    const uint4 m = *reinterpret_cast<const uint4 *>(mrow + 4 * static_cast<size_t>(x));
    const uint2 f = *reinterpret_cast<const uint2 *>(frow + 2 * static_cast<size_t>(x));
    xy[0] = m.x; xy[1] = m.y; xy[2] = m.z; xy[3] = m.w;
    fr[0] = f.x & 0xFFFFu; fr[1] = f.x >> 16; fr[2] = f.y & 0xFFFFu; fr[3] = f.y >> 16;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int xx = x + (k < n ? k : 0);
      const uint16_t *m = reinterpret_cast<const uint16_t *>(mrow + 4 * static_cast<size_t>(xx));
      xy[k] = static_cast<uint32_t>(m[0]) | (static_cast<uint32_t>(m[1]) << 16);
      fr[k] = *reinterpret_cast<const uint16_t *>(frow + 2 * static_cast<size_t>(xx));
    }
  }
}

// n pixels (C channels each) of a destination row from x on.
template <int C>
__device__ __forceinline__ void store_s16_px(const RemapS16Args &a, int x, int y, int n,
                                             const uint32_t (&out)[4]) {
  uint8_t *d = a.dst + static_cast<size_t>(y) * a.dst_stride + static_cast<size_t>(x) * C;
  if (C == 1 && n == 4 && (reinterpret_cast<uintptr_t>(d) & 3) == 0) {
    store_nt(reinterpret_cast<uint32_t *>(d), out[0] | (out[1] << 8) | (out[2] << 16) | (out[3] << 24));
  } else if (C == 4 && n == 4 && (reinterpret_cast<uintptr_t>(d) & 15) == 0) {
    store_nt(reinterpret_cast<uint4 *>(d), make_uint4(out[0], out[1], out[2], out[3]));
  } else {
    const bool al = (reinterpret_cast<uintptr_t>(d) & 3) == 0;
    for (int k = 0; k < n; ++k) store_px<C>(d + k * C, out[k], al);
  }
}

// Destination tile (bx, by) of gather_tile_w(C) x gather_tile_h(C) pixels.
template <int C, int BORDER>
__device__ __forceinline__ void remap_s16point5_tile(const RemapS16Args &a, int bx, int by) {
  constexpr int TW = gather_tile_w(C), TH = gather_tile_h(C);
  const int q = threadIdx.x % (TW / 4), r = threadIdx.x / (TW / 4);
  const int x = bx * TW + 4 * q, y = by * TH + r;
  if (x >= a.dst_width || y >= a.dst_height) return;
  const int n = min(4, a.dst_width - x);
  uint32_t bv = 0;
#pragma unroll
  for (int c = 0; c < C; ++c)
    bv |= static_cast<uint32_t>(a.border_value[C == 4 ? c : 0]) << (8 * c);
  uint32_t xy[4], fr[4];
  load_s16_maps(a, x, y, n, xy, fr);
  const int w = a.src_width, h = a.src_height;
  uint32_t out[4];
  if (w >= 2) {
    const uint32_t span = static_cast<uint32_t>(h - 1) * static_cast<uint32_t>(a.src_stride) +
                          static_cast<uint32_t>(w * C);
    const SrcWindow src = make_window(a.src, span);
    const uint32_t stride = static_cast<uint32_t>(a.src_stride);
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = remap_s16_px<C, BORDER>(src, stride, w, h, xy[k], fr[k], bv);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = remap_s16_px_narrow<C, BORDER>(a, xy[k], fr[k], bv);
  }
  store_s16_px<C>(a, x, y, n, out);
}

// ---------------------------------------------------------------------------
// Float gathers (transform_neon.h:189-491).
//
// Coordinate conversions.  The reference converts with FCVTM* (floor) and
// FCVTA* (round half away from zero), saturating, NaN -> 0, then clamps the
// unsigned results to the image (replicate) or range-tests the signed ones
// against it (constant).  Source sizes are below 2^24, so saturating at
// +-2^24 instead of at the integer limits gives the same clamps and tests;
// written this way the conversions are branch-free.
constexpr float kIndexLimit = 16777216.0f;  // 2^24

__device__ __forceinline__ uint32_t floor_u(float f) {  // fmaxf maps NaN to 0
  return static_cast<uint32_t>(fminf(fmaxf(floorf(f), 0.0f), kIndexLimit));
}

__device__ __forceinline__ int32_t floor_s(float f) {
  const float r = f == f ? floorf(f) : 0.0f;
  return static_cast<int32_t>(fminf(fmaxf(r, -kIndexLimit), kIndexLimit));
}

__device__ __forceinline__ uint32_t round_u(float f) {  // roundf: ties away from zero
  return static_cast<uint32_t>(fminf(fmaxf(roundf(f), 0.0f), kIndexLimit));
}

__device__ __forceinline__ int32_t round_s(float f) {
  const float r = f == f ? roundf(f) : 0.0f;
  return static_cast<int32_t>(fminf(fmaxf(r, -kIndexLimit), kIndexLimit));
}

// 1/w rounded to nearest even.  The hardware reciprocal refined by one
// fused Newton step is bit-identical to the IEEE quotient 1.0f / w for every
// |w| in [2^-125, 2^125] (kcv_selfcheck_reciprocal compares the two over all
// floats); a wave with any lane outside that range divides instead.
__device__ __forceinline__ bool recip_refined_ok(float w) {
  const float aw = fabsf(w);
  return aw >= 0x1p-125f && aw <= 0x1p125f;
}

__device__ __forceinline__ float recip_refined(float w) {
  const float y0 = __builtin_amdgcn_rcpf(w);
  return fmaf(fmaf(-w, y0, 1.0f), y0, y0);
}

// Source image of a float gather.  When the image spans less than
// kBufferSpanLimit bytes, its bytes are read through the
// aligned-window buffer resource at 32-bit offsets (no 64-bit address
// arithmetic per tap), else with global loads at 64-bit addresses, which
// may be unaligned (gfx950 global accesses are byte-addressed in unaligned
// mode) and never read past the last byte of the image; the choice is
// uniform over a launch.
struct Src {
  const uint8_t *p;
  size_t stride;
  uint32_t stride32;
  uint32_t xmax, ymax;
  uint32_t bv;  // border value bytes, one per channel
  SrcWindow win;  // meaningful when src_fits_buffer
};

template <bool BUF, int NB>
__device__ __forceinline__ uint32_t fetch(const Src &s, uint32_t y, uint32_t xb) {
  if (BUF) {
    const int off = static_cast<int>(y * s.stride32 + xb + s.win.lead);
    if (NB == 1) return __builtin_amdgcn_raw_buffer_load_b8(s.win.rsrc, off, 0, 0);
    if (NB == 2) return __builtin_amdgcn_raw_buffer_load_b16(s.win.rsrc, off, 0, 0);
    return __builtin_amdgcn_raw_buffer_load_b32(s.win.rsrc, off, 0, 0);
  }
  const uint8_t *q = s.p + static_cast<size_t>(y) * s.stride + xb;
  typedef uint16_t u16u __attribute__((aligned(1)));
  typedef uint32_t u32u __attribute__((aligned(1)));
  if (NB == 1) return *q;
  if (NB == 2) return *reinterpret_cast<const u16u *>(q);
  return *reinterpret_cast<const u32u *>(q);
}

template <bool BUF>
__device__ __forceinline__ uint64_t fetch64(const Src &s, uint32_t y, uint32_t xb) {
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
  if (BUF) {
    const int off = static_cast<int>(y * s.stride32 + xb + s.win.lead);
    const u32x2 v = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(s.win.rsrc, off, 0, 0));
    return (static_cast<uint64_t>(v.y) << 32) | v.x;
  }
  typedef uint64_t u64u __attribute__((aligned(1)));
  return *reinterpret_cast<const u64u *>(s.p + static_cast<size_t>(y) * s.stride + xb);
}

// Pixel x of row y, C bytes packed in a dword.
template <int C, bool BUF>
__device__ __forceinline__ uint32_t fetch_px(const Src &s, uint32_t y, uint32_t x) {
  const uint32_t xb = x * C;
  if (C == 1) return fetch<BUF, 1>(s, y, xb);
  if (C == 2) return fetch<BUF, 2>(s, y, xb);
  if (C == 3) {
    if (BUF) return window4(s.win, y * s.stride32 + xb) & 0xFFFFFFu;
    return fetch<BUF, 2>(s, y, xb) | (fetch<BUF, 1>(s, y, xb + 2) << 16);
  }
  return fetch<BUF, 4>(s, y, xb);
}

// Pixels cx and cx + 1 of row y (cx + 1 <= xmax), or pixel 0 twice for an
// image one pixel wide.  Through the buffer resource, 1-3 channel pairs are
// cut out of one aligned window; 4-channel pairs are one 8-byte load at the
// pixel, dword aligned when the image and its stride are.
template <int C, bool BUF>
__device__ __forceinline__ void fetch_pair(const Src &s, uint32_t y, uint32_t cx, uint32_t &p0,
                                           uint32_t &p1) {
  if (s.xmax == 0) {
    p0 = p1 = fetch_px<C, BUF>(s, y, 0);
    return;
  }
  const uint32_t xb = cx * C;
  if (BUF && C < 4) {
    const uint32_t off = y * s.stride32 + xb;
    if (C == 3) {
      uint32_t lo, hi;
      window6(s.win, off, lo, hi);
      p0 = lo & 0xFFFFFFu;
      p1 = (lo >> 24) | (hi << 8);
    } else {
      const uint32_t t = window4(s.win, off);
      p0 = C == 1 ? t & 0xFFu : t & 0xFFFFu;
      p1 = C == 1 ? (t >> 8) & 0xFFu : t >> 16;
    }
    return;
  }
  if (C == 1) {
    const uint32_t t = fetch<BUF, 2>(s, y, xb);
    p0 = t & 0xFFu;
    p1 = t >> 8;
  } else if (C == 2) {
    const uint32_t t = fetch<BUF, 4>(s, y, xb);
    p0 = t & 0xFFFFu;
    p1 = t >> 16;
  } else if (C == 3) {
    const uint32_t lo = fetch<BUF, 4>(s, y, xb), hi = fetch<BUF, 2>(s, y, xb + 4);
    p0 = lo & 0xFFFFFFu;
    p1 = (lo >> 24) | (hi << 8);
  } else {
    const uint64_t t = fetch64<BUF>(s, y, xb);
    p0 = static_cast<uint32_t>(t);
    p1 = static_cast<uint32_t>(t >> 32);
  }
}

// Gathered corners of one pixel, ready for the lerp.
struct Quad {
  uint32_t a, b, c, d;
  float xfrac, yfrac;
};

// load_quad_pixels_{replicate,constant} (transform_neon.h:189-329).  Each
// row is one pair load at cx = the left tap clamped to [0, xmax - 1]; the
// taps are then picked from the pair (index 0 or 1) and, for the constant
// border, replaced by the border value when outside the image.  No branch
// depends on the pixel, so the loads of all pixels of a thread overlap.
template <int C, int BORDER, bool BUF>
__device__ __forceinline__ Quad gather_linear(const Src &s, float xf, float yf) {
  Quad q;
  q.xfrac = xf - floorf(xf);
  q.yfrac = yf - floorf(yf);
  uint32_t cx, y0, y1, i0, i1;
  bool ina = true, inb = true, inc = true, ind = true;
  if (BORDER == kReplicate) {
    const uint32_t x0 = min(floor_u(xf), s.xmax);
    y0 = min(floor_u(yf), s.ymax);
    const uint32_t x1 = x0 + ((xf >= 0.0f && x0 < s.xmax) ? 1u : 0u);
    y1 = y0 + ((yf >= 0.0f && y0 < s.ymax) ? 1u : 0u);
    cx = s.xmax ? min(x0, s.xmax - 1) : 0u;
    i0 = x0 - cx;
    i1 = x1 - cx;
  } else {
    const int32_t ix = floor_s(xf), iy = floor_s(yf);
    const uint32_t x0 = static_cast<uint32_t>(ix), x1 = x0 + 1u;
    const uint32_t ya = static_cast<uint32_t>(iy), yb = ya + 1u;
    const bool x0in = x0 <= s.xmax, x1in = x1 <= s.xmax;
    const bool y0in = ya <= s.ymax, y1in = yb <= s.ymax;
    ina = x0in && y0in;
    inb = x1in && y0in;
    inc = x0in && y1in;
    ind = x1in && y1in;
    // rows clamped into the image: an inside tap reads its own row, an
    // outside one a row next to its neighbours' (its value is replaced)
    const int32_t ymax = static_cast<int32_t>(s.ymax), ty = min(iy, ymax);
    y0 = static_cast<uint32_t>(max(ty, 0));
    y1 = static_cast<uint32_t>(min(max(ty + 1, 0), ymax));
    cx = s.xmax ? static_cast<uint32_t>(max(min(ix, static_cast<int32_t>(s.xmax) - 1), 0)) : 0u;
    i0 = (x0 - cx) & 1u;  // meaningful only when the tap is inside
    i1 = (x1 - cx) & 1u;
  }
  uint32_t a0, a1, b0, b1;
  fetch_pair<C, BUF>(s, y0, cx, a0, a1);
  fetch_pair<C, BUF>(s, y1, cx, b0, b1);
  q.a = i0 ? a1 : a0;
  q.b = i1 ? a1 : a0;
  q.c = i0 ? b1 : b0;
  q.d = i1 ? b1 : b0;
  if (BORDER != kReplicate) {
    q.a = ina ? q.a : s.bv;
    q.b = inb ? q.b : s.bv;
    q.c = inc ? q.c : s.bv;
    q.d = ind ? q.d : s.bv;
  }
  return q;
}

// FCVTAU of a lerp result: round half away from zero, NaN to 0.  A lerp of
// bytes with a fraction in [0, 1) is NaN or lies in [0, 255] (the fused
// steps cannot round outside the range of their inputs), so no saturation
// test is needed; fmaxf(v, 0) maps a (quiet) NaN to 0.  v_cvt_rpi_i32_f32
// (floor(x + 0.5) without an intermediate rounding) equals round-half-away
// for every float in [0, 256] (kcv_selfcheck_rounding compares the two).
__device__ __forceinline__ uint32_t round_half_away_u8(float c) {
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(c));
  return static_cast<uint32_t>(r);
}

__device__ __forceinline__ uint32_t round_lerp_u8(float v) {
  return round_half_away_u8(fmaxf(v, 0.0f));
}

typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));

// Channels 2k and 2k + 1 of a pixel as two f16 values 1024 + byte: the byte
// goes under an exponent byte 0x64, and f16 spacing is 1 on [1024, 2048), so
// both are exact.
__device__ __forceinline__ h16x2 biased_pair(uint32_t v, int k) {
  return __builtin_bit_cast(h16x2, __builtin_amdgcn_perm(0x64646464u, v,
                                                         k == 0 ? 0x04010400u : 0x04030402u));
}

// lerp_2d (transform_neon.h:331-337) per channel.  FINITE: the fractions
// are known finite (then the result lies in [0, 255] and needs no NaN
// mapping).
//
// Channels are taken in pairs: one packed f16 subtraction gives two exact tap
// differences (|b - a| <= 255), another two exact tap values (1024 + a -
// 1024), and each horizontal fma takes them as f16 operands (v_fma_mix_f32:
// f16 inputs widened exactly, one rounding), so it equals the float fma of
// the same values: 40 instead of 48 VALU instructions per 4-channel pixel.
// An odd last channel takes float taps directly.
template <int C, bool FINITE = false>
__device__ __forceinline__ uint32_t lerp_quad(const Quad &q) {
  uint32_t out = 0;
  const h16x2 bias = {static_cast<_Float16>(1024.0f), static_cast<_Float16>(1024.0f)};
#pragma unroll
  for (int k = 0; k < C / 2; ++k) {
    const h16x2 a = biased_pair(q.a, k), b = biased_pair(q.b, k);
    const h16x2 c = biased_pair(q.c, k), d = biased_pair(q.d, k);
    const h16x2 dab = b - a, dcd = d - c;
    h16x2 pa = a - bias, pc = c - bias;
    // keeps each pair in one register, so one packed subtraction makes it
    // (else each half gets an f16 subtraction of its own)
    asm("" : "+v"(pa), "+v"(pc));
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const float line0 = fmaf(static_cast<float>(dab[j]), q.xfrac, static_cast<float>(pa[j]));
      const float line1 = fmaf(static_cast<float>(dcd[j]), q.xfrac, static_cast<float>(pc[j]));
      const float res = fmaf(line1 - line0, q.yfrac, line0);
      out |= (FINITE ? round_half_away_u8(res) : round_lerp_u8(res)) << (8 * (2 * k + j));
    }
  }
  if constexpr (C % 2 == 1) {
    constexpr int c = C - 1;
    const float pa = static_cast<float>(byte_of(q.a, c));
    const float pb = static_cast<float>(byte_of(q.b, c));
    const float pc = static_cast<float>(byte_of(q.c, c));
    const float pd = static_cast<float>(byte_of(q.d, c));
    const float line0 = fmaf(pb - pa, q.xfrac, pa);
    const float line1 = fmaf(pd - pc, q.xfrac, pc);
    const float res = fmaf(line1 - line0, q.yfrac, line0);
    out |= (FINITE ? round_half_away_u8(res) : round_lerp_u8(res)) << (8 * c);
  }
  return out;
}

// transform_pixels_{replicate,constant} (transform_neon.h:339-491).
template <int C, int BORDER, bool BUF>
__device__ __forceinline__ uint32_t gather_nearest(const Src &s, float xf, float yf) {
  if (BORDER == kReplicate) {
    const uint32_t x = min(round_u(xf), s.xmax);
    const uint32_t y = min(round_u(yf), s.ymax);
    return fetch_px<C, BUF>(s, y, x);
  }
  const uint32_t x = static_cast<uint32_t>(round_s(xf));
  const uint32_t y = static_cast<uint32_t>(round_s(yf));
  const bool in = x <= s.xmax && y <= s.ymax;
  const uint32_t v = fetch_px<C, BUF>(s, in ? y : 0u, in ? x : 0u);
  return in ? v : s.bv;
}

template <int C>
__device__ __forceinline__ Src make_src(const uint8_t *p, size_t stride, int w, int h,
                                        const uint8_t *border_value) {
  Src s;
  s.p = p;
  s.stride = stride;
  s.stride32 = static_cast<uint32_t>(stride);
  s.xmax = static_cast<uint32_t>(w - 1);
  s.ymax = static_cast<uint32_t>(h - 1);
  s.bv = 0;
#pragma unroll
  for (int c = 0; c < C; ++c) s.bv |= static_cast<uint32_t>(border_value[c]) << (8 * c);
  const uint64_t span = static_cast<uint64_t>(h - 1) * stride + static_cast<uint64_t>(w) * C;
  s.win = make_window(p, span < kBufferSpanLimit ? span : 0);
  return s;
}

// Whether the gathers of `s` go through its buffer resource: the image
// spans less than kBufferSpanLimit bytes.
__device__ __forceinline__ bool src_fits_buffer(const Src &s, int c) {
  const uint64_t span =
      static_cast<uint64_t>(s.ymax) * s.stride + (static_cast<uint64_t>(s.xmax) + 1) * c;
  return span < kBufferSpanLimit;
}

// The 4 destination pixels (x .. x + n - 1) of this thread, C bytes each,
// as one 4/8/12/16-byte store when whole and aligned.
template <int C>
__device__ __forceinline__ void store_quad_px(uint8_t *d, int n, const uint32_t (&o)[4]) {
  const uintptr_t al = reinterpret_cast<uintptr_t>(d);
  if (n == 4) {
    if (C == 1 && (al & 3) == 0) {
      store_nt(reinterpret_cast<uint32_t *>(d), o[0] | (o[1] << 8) | (o[2] << 16) | (o[3] << 24));
      return;
    }
    if (C == 2 && (al & 7) == 0) {
      store_nt(reinterpret_cast<uint2 *>(d), make_uint2(o[0] | (o[1] << 16), o[2] | (o[3] << 16)));
      return;
    }
    if (C == 3 && (al & 3) == 0) {
      uint32_t *w = reinterpret_cast<uint32_t *>(d);
      store_nt(w, o[0] | (o[1] << 24));
      store_nt(w + 1, (o[1] >> 8) | (o[2] << 16));
      store_nt(w + 2, (o[2] >> 16) | (o[3] << 8));
      return;
    }
    if (C == 4 && (al & 15) == 0) {
      store_nt(reinterpret_cast<uint4 *>(d), make_uint4(o[0], o[1], o[2], o[3]));
      return;
    }
  }
  for (int k = 0; k < n; ++k) store_px<C>(d + k * C, o[k], (al & 3) == 0 && C == 4);
}

// Shared body of the float gathers over a gather_tile_w(C) x gather_tile_h(C) tile:
// a thread owns 4 horizontally adjacent pixels of one row; `coords(y, x, n,
// xf, yf)` fills the source positions of those pixels.
//
// Linear gathers of a thread whose four pixels all have 0 <= x < xmax and
// 0 <= y < ymax (so every tap is inside the image, and NaNs are excluded)
// take a short path: under either border rule the pair starts at floor(x)
// with taps 0 and 1, the rows are floor(y) and floor(y) + 1, and no border
// value or clamp applies.
template <int C, int INTERP, int BORDER, bool BUF>
__device__ __forceinline__ void gather_quad(const Src &s, const float (&xf)[4],
                                            const float (&yf)[4], uint32_t (&out)[4]) {
  if (INTERP == 1) {
    const float xlim = static_cast<float>(s.xmax), ylim = static_cast<float>(s.ymax);
    bool inner = true;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      inner &= (xf[k] >= 0.0f) & (xf[k] < xlim) & (yf[k] >= 0.0f) & (yf[k] < ylim);
    Quad q[4];
    if (inner) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float fx = floorf(xf[k]), fy = floorf(yf[k]);
        const uint32_t x0 = static_cast<uint32_t>(fx), y0 = static_cast<uint32_t>(fy);
        fetch_pair<C, BUF>(s, y0, x0, q[k].a, q[k].b);
        fetch_pair<C, BUF>(s, y0 + 1, x0, q[k].c, q[k].d);
        q[k].xfrac = xf[k] - fx;
        q[k].yfrac = yf[k] - fy;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) out[k] = lerp_quad<C, true>(q[k]);
      return;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = gather_linear<C, BORDER, BUF>(s, xf[k], yf[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = lerp_quad<C>(q[k]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = gather_nearest<C, BORDER, BUF>(s, xf[k], yf[k]);
  }
}

template <int C, int INTERP, int BORDER, class Coords>
__device__ __forceinline__ void gather_tiles(const Src &s, uint8_t *dst, size_t dst_stride,
                                             int dst_width, int dst_height, int bx,
                                             int by, const Coords &coords) {
  constexpr int TW = gather_tile_w(C), TH = gather_tile_h(C);
  const int q = threadIdx.x % (TW / 4), r = threadIdx.x / (TW / 4);
  const int x = bx * TW + 4 * q, y = by * TH + r;
  if (x >= dst_width || y >= dst_height) return;
  const int n = min(4, dst_width - x);
  float xf[4], yf[4];
  coords(y, x, n, xf, yf);
  uint32_t out[4];
  if (src_fits_buffer(s, C))
    gather_quad<C, INTERP, BORDER, true>(s, xf, yf, out);
  else
    gather_quad<C, INTERP, BORDER, false>(s, xf, yf, out);
  store_quad_px<C>(dst + static_cast<size_t>(y) * dst_stride + static_cast<size_t>(x) * C, n,
                   out);
}

template <int C, int INTERP, int BORDER>
__device__ __forceinline__ void remap_f32_tile(const RemapF32Args &a, int bx, int by) {
  const Src s = make_src<C>(a.src, a.src_stride, a.src_width, a.src_height, a.border_value);
  auto coords = [&](int y, int x, int n, float (&xf)[4], float (&yf)[4]) {
    const uint8_t *mx = reinterpret_cast<const uint8_t *>(a.mapx) + static_cast<size_t>(y) * a.mapx_stride;
    const uint8_t *my = reinterpret_cast<const uint8_t *>(a.mapy) + static_cast<size_t>(y) * a.mapy_stride;
    const float *px = reinterpret_cast<const float *>(mx) + x;
    const float *py = reinterpret_cast<const float *>(my) + x;
    if (n == 4 && ((reinterpret_cast<uintptr_t>(px) | reinterpret_cast<uintptr_t>(py)) & 15) == 0) {
      const float4 vx = *reinterpret_cast<const float4 *>(px);
      const float4 vy = *reinterpret_cast<const float4 *>(py);
      xf[0] = vx.x; xf[1] = vx.y; xf[2] = vx.z; xf[3] = vx.w;
      yf[0] = vy.x; yf[1] = vy.y; yf[2] = vy.z; yf[3] = vy.w;
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int kk = k < n ? k : 0;
        xf[k] = px[kk];
        yf[k] = py[kk];
      }
    }
  };
  gather_tiles<C, INTERP, BORDER>(s, a.dst, a.dst_stride, a.dst_width, a.dst_height, bx, by,
                                  coords);
}

// warp_perspective_neon.cpp:55-76 and :98-102: the row terms are shared by
// the pixels of a row, each pixel adds its column term with one fma.
struct WarpRow {
  float tx0, ty0, tw0;
};

__device__ __forceinline__ WarpRow warp_row(const WarpArgs &a, int y) {
  const float fy = static_cast<float>(y);
  return {fmaf(a.t[1], fy, a.t[2]), fmaf(a.t[4], fy, a.t[5]), fmaf(a.t[7], fy, a.t[8])};
}

template <int C, int INTERP, int BORDER>
__device__ __forceinline__ void warp_tile(const WarpArgs &a, int bx, int by) {
  const Src s = make_src<C>(a.src, a.src_stride, a.src_width, a.src_height, a.border_value);
  auto coords = [&](int y, int x, int, float (&xf)[4], float (&yf)[4]) {
    const WarpRow row = warp_row(a, y);
    float tx[4], ty[4], tw[4];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float fx = static_cast<float>(static_cast<uint32_t>(x + k));
      tx[k] = fmaf(fx, a.t[0], row.tx0);
      ty[k] = fmaf(fx, a.t[3], row.ty0);
      tw[k] = fmaf(fx, a.t[6], row.tw0);
      ok &= recip_refined_ok(tw[k]);
    }
    // 1/w for the thread's four weights with one vote of the wave: the
    // refined reciprocal when every lane's weights allow it (it equals the
    // IEEE quotient there), else the division
    float r[4];
    if (__builtin_amdgcn_ballot_w64(!ok) == 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) r[k] = recip_refined(tw[k]);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) r[k] = 1.0f / tw[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float iw = (a.check_zero && tw[k] == 0.0f) ? 0.0f : r[k];
      xf[k] = tx[k] * iw;
      yf[k] = ty[k] * iw;
    }
  };
  gather_tiles<C, INTERP, BORDER>(s, a.dst, a.dst_stride, a.dst_width, a.dst_height, bx, by,
                                  coords);
}

}  // namespace xform
}  // namespace kcv

#endif  // KCV_TRANSFORM_BODY_H
