// Device bodies and host helpers of the linear resize (per-image kernels in
// resize.hip, batched kernels in batch.hip).
//
// Path selection follows src/resize/resize_linear_api.cpp:82-216:
//   * exact 0.5 x 0.5            -> quarter kernel, (a+b+c+d+2) >> 2
//                                   (src/resize/resize_to_quarter_neon.cpp:179-213)
//   * 1 channel, exact 2x or 4x  -> closed-form upsizer
//                                   (src/resize/resize_linear_neon.cpp:61-...)
//   * everything else            -> generic fixed-point lerp
//                                   (src/resize/resize_linear_generic_u8_neon.h)
//
// The generic path reproduces the reference's column coordinates exactly,
// including the accumulated 16.16 step with its periodic recalibration, the
// pull-backs at the right edge and the table-lookup semantics of the vector
// code (an index beyond the loaded bytes reads 0).  A small kernel evaluates
// that generator in closed form per destination element (every lane of every
// vector is a pure function of its vector index) and writes a column table
// and a row table to the workspace; the main kernel then gathers and does the
// two 8-bit lerps, vertical first.

#ifndef KCV_RESIZE_BODY_H
#define KCV_RESIZE_BODY_H

#include "kcv_internal.h"

namespace kcv {
namespace rsz {


constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// Host/device-shared coordinate arithmetic
// (resize_linear_generic_u8_neon.h:73-88).
__host__ __device__ __forceinline__ int64_t rounding_div(uint64_t nom,
                                                         uint64_t denom) {
  return static_cast<int64_t>((nom + denom / 2) / denom);
}

__host__ __device__ __forceinline__ int64_t aligned_scale(uint64_t x,
                                                          uint64_t nom,
                                                          uint64_t denom) {
  return rounding_div(((x << 16) + 32768) * nom, denom) - 32768 + 128;
}

struct GenericPlan {
  int src_width, dst_width, channels;
  int ratio;           // source vectors loaded per 16 destination bytes
  int upsize;          // horizontal upsizing variant
  int set_rightmost;   // 3ch r3 with inverse scale >= 2.8
  int64_t two_x;       // pairs of full vectors
  int64_t half_n;      // half vectors
  int full_read;       // bytes loaded per full vector
  int half_read;       // bytes loaded per half vector
  int dst_height, src_height;
};

inline GenericPlan make_plan(const ResizeArgs &a) {
  GenericPlan p{};
  const int64_t sw = a.src_width, dw = a.dst_width, C = a.channels;
  p.src_width = a.src_width;
  p.dst_width = a.dst_width;
  p.channels = a.channels;
  p.src_height = a.src_height;
  p.dst_height = a.dst_height;
  auto set = [&](int r, int up) {
    p.ratio = r;
    p.upsize = up;
  };
  if (C == 1) {
    if (dw * 14 > sw * 15) set(1, 1);
    else if (dw >= sw) set(2, 1);
    else if (dw * 2 >= sw) set(2, 0);
    else set(3, 0);
  } else if (C == 2) {
    if (dw * 6 > sw * 7) set(1, 1);
    else if (dw >= sw) set(2, 1);
    else if (dw * 2 >= sw) set(2, 0);
    else set(3, 0);
  } else if (C == 3) {
    if (dw * 2 > sw * 3) set(1, 1);
    else if (dw >= sw) set(2, 1);
    else {
      const double inv = static_cast<double>(sw) / static_cast<double>(dw);
      if (inv < 1.8) set(2, 0);
      else {
        set(3, 0);
        p.set_rightmost = inv >= 2.8;  // resize_linear_generic_neon.cpp:39-52
      }
    }
  } else {
    if (dw * 2 > sw * 3) set(1, 1);
    else if (dw >= sw) set(2, 1);
    else if (dw * 2 >= sw) set(2, 0);
    else set(3, 0);
  }
  p.full_read = 16 * p.ratio;
  if (C == 3)
    p.half_read = p.upsize ? 16 : 32;
  else
    p.half_read = (p.ratio == 3 && !p.upsize) ? 32 : 16;
  // calculate_num_of_vector_paths (resize_linear_generic_u8_neon.h:151-160)
  p.two_x = (sw * C >= p.full_read) ? (dw * C) / 32 : 0;
  const int64_t remaining = dw * C - p.two_x * 32;
  p.half_n = (remaining + 7) / 8;
  return p;
}

struct Tap {
  int64_t base;
  int idx0, idx1;  // tbl indices as unsigned bytes
  int frac;
  int load;
  bool direct_right;
};

__device__ __forceinline__ int to_u8(int v) { return v & 0xFF; }
__device__ __forceinline__ int to_s8(int v) {
  v &= 0xFF;
  return v >= 128 ? v - 256 : v;
}

__device__ __forceinline__ uint32_t scale_x(const GenericPlan &p, int64_t dx) {
  return static_cast<uint32_t>(
      rounding_div((static_cast<uint64_t>(dx) * p.src_width) << 16, p.dst_width));
}

// One lane of a full or half vector for 1, 2 or 4 channels
// (resize_linear_generic_u8_neon.h:301-401).
__device__ Tap lane_124(const GenericPlan &p, int64_t sx_base, int64_t sx_fixp,
                        int lane, int load) {
  const int C = p.channels;
  const int32_t xfrac0 = static_cast<int32_t>(sx_fixp - (sx_base << 16));
  const uint32_t delta = scale_x(p, lane / C) + static_cast<uint32_t>(xfrac0);
  int i0 = to_s8(static_cast<int>(delta >> 16));
  int i1 = to_s8(i0 + 1);
  if (p.upsize) {
    const int mx = (load - 1) / C;
    i0 = min(max(i0, 0), mx);
    i1 = min(max(i1, 0), mx);
  }
  int u0 = to_u8(i0), u1 = to_u8(i1);
  if (C > 1) {
    const int sh = C / 2;
    const int off = lane % C;
    u0 = to_u8((u0 << sh) + off);
    u1 = to_u8((u1 << sh) + off);
  }
  return Tap{sx_base * C, u0, u1, static_cast<int>((delta >> 8) & 0xFF), load,
             false};
}

__device__ Tap full_124(const GenericPlan &p, int64_t v, int lane) {
  const int C = p.channels;
  const int64_t i = v >> 1;
  const int64_t i0 = i & ~int64_t{255};
  const int64_t dx0 = i0 * 2 * (16 / C);
  const int64_t step =
      rounding_div((static_cast<uint64_t>(p.src_width) * 16 / C) << 16,
                   p.dst_width);
  const int64_t sx_fixp =
      aligned_scale(dx0, p.src_width, p.dst_width) + (v - 2 * i0) * step;
  const int64_t max_sx =
      (static_cast<int64_t>(p.src_width) * C - p.full_read) / C;
  const int64_t base = min(max(sx_fixp >> 16, int64_t{0}), max_sx);
  return lane_124(p, base, sx_fixp, lane, p.full_read);
}

__device__ Tap half_124(const GenericPlan &p, int64_t h, int lane) {
  const int C = p.channels;
  const int64_t d0 = p.two_x * 32 / C;  // first half vector's pixel
  const int64_t step_px = 8 / C;
  const int64_t dx = d0 + h * step_px;
  const int64_t dxp = min(dx, static_cast<int64_t>(p.dst_width) - step_px);
  const int64_t one_px =
      rounding_div(static_cast<uint64_t>(p.src_width) << 16, p.dst_width);
  const int64_t half_step =
      rounding_div((static_cast<uint64_t>(p.src_width) * 8 / C) << 16,
                   p.dst_width);
  const int64_t sx_fixp = aligned_scale(d0, p.src_width, p.dst_width) +
                          h * half_step - (dx - dxp) * one_px;
  const int64_t max_sx =
      (static_cast<int64_t>(p.src_width) * C - p.half_read) / C;
  const int64_t base = min(max(sx_fixp >> 16, int64_t{0}), max_sx);
  return lane_124(p, base, sx_fixp, lane, p.half_read);
}

// 3-channel full vectors (resize_linear_generic_u8_neon.h:404-628): lanes
// of the r, g and b vectors of a 16-pixel triplet.
__device__ Tap full_3(const GenericPlan &p, int64_t v, int lane) {
  const int64_t t = v / 3;
  const int phase = static_cast<int>(v - 3 * t);
  int64_t tr = 0;
  if (t >= 512) tr = 512 + 513 * ((t - 512) / 513);
  const int64_t triplet_step =
      rounding_div(static_cast<uint64_t>(p.src_width) << 20, p.dst_width);
  const int64_t sx_fixp =
      aligned_scale(16 * tr, p.src_width, p.dst_width) + (t - tr) * triplet_step;
  const int64_t sx_base = sx_fixp >> 16;
  const int64_t src_base = sx_base * 3;
  const int32_t xfrac0 = static_cast<int32_t>(sx_fixp & 0xFFFF);
  const int64_t sw3 = static_cast<int64_t>(p.src_width) * 3;
  const int64_t max_base = sw3 > p.full_read ? sw3 - p.full_read : 0;
  // lane -> pixel offset within the triplet and channel offset
  auto px_of = [&](int l) { return (16 * phase + l) / 3; };
  auto ch_of = [&](int l) { return (16 * phase + l) % 3; };
  auto idx0_of = [&](int l) {
    const uint32_t delta = scale_x(p, px_of(l)) + static_cast<uint32_t>(xfrac0);
    int i0 = to_s8(static_cast<int>(delta >> 16));
    if (p.upsize) {
      const int mn = static_cast<int>(min(max(-sx_base, int64_t{-128}), int64_t{127}));
      const uint64_t room = static_cast<uint64_t>(p.src_width) - 1 -
                            static_cast<uint64_t>(sx_base);
      const int mx = static_cast<int>(room > 127 ? 127 : room);
      i0 = min(max(i0, mn), mx);
    }
    return to_s8(to_s8(i0 * 3) + ch_of(l));
  };
  const uint32_t delta =
      scale_x(p, px_of(lane)) + static_cast<uint32_t>(xfrac0);
  int i0 = to_s8(static_cast<int>(delta >> 16));
  int i1 = to_s8(i0 + 1);
  if (p.upsize) {
    const int mn = static_cast<int>(min(max(-sx_base, int64_t{-128}), int64_t{127}));
    const uint64_t room =
        static_cast<uint64_t>(p.src_width) - 1 - static_cast<uint64_t>(sx_base);
    const int mx = static_cast<int>(room > 127 ? 127 : room);
    i0 = min(max(i0, mn), mx);
    i1 = min(max(i1, mn), mx);
  }
  i0 = to_s8(to_s8(i0 * 3) + ch_of(lane));
  i1 = to_s8(to_s8(i1 * 3) + ch_of(lane));
  const int mn_idx = min(min(idx0_of(0), idx0_of(1)), idx0_of(2));
  const int64_t final_base =
      min(max(src_base + mn_idx, int64_t{0}), max_base);
  const int adj = to_s8(static_cast<int>(final_base - src_base));
  return Tap{final_base, to_u8(i0 - adj), to_u8(i1 - adj),
             static_cast<int>((delta >> 8) & 0xFF), p.full_read,
             p.set_rightmost && lane == 15};
}

// 3-channel half vectors, filled scalarly (:630-703).
__device__ Tap half_3(const GenericPlan &p, int64_t start, int lane) {
  const int64_t sw = p.src_width;
  const int64_t sw3 = sw * 3;
  const int64_t max_base = sw3 > p.half_read ? sw3 - p.half_read : 0;
  const int64_t dx = start / 3;
  const int inpx = static_cast<int>(start % 3);
  const int64_t sxp0 = aligned_scale(dx, p.src_width, p.dst_width);
  const int64_t sxp1 = aligned_scale(dx + 1, p.src_width, p.dst_width);
  const int64_t sei0 = ((sxp0 >> 16) * 3) + inpx;
  const int64_t sei1 = (sxp1 >> 16) * 3;
  const int64_t base = min(max(min(sei0, sei1), int64_t{0}), max_base);
  const int64_t one_px =
      rounding_div(static_cast<uint64_t>(p.src_width) << 16, p.dst_width);
  // Which pixel of the half vector this lane belongs to, and its channel.
  int pix, k;
  if (lane < 3 - inpx) {
    pix = 0;
    k = inpx + lane;
  } else {
    const int l = lane - (3 - inpx);
    pix = 1 + l / 3;
    k = l % 3;
  }
  const int64_t sx = sxp0 + pix * one_px;
  int64_t s0 = sx >> 16, s1 = s0 + 1;
  if (p.upsize) {
    s0 = min(max(s0, int64_t{0}), sw - 1);
    s1 = min(max(s1, int64_t{0}), sw - 1);
  }
  int frac = static_cast<int>((sx & 0xFFFF) >> 8);
  // The first pixel tests against the base, the second does not, later
  // ones do again (fill_half_constants_scalarly, :654-702).
  const int64_t sei = pix == 0 ? sei0 : (sx >> 16) * 3;
  if (pix != 1 && sei < base) frac = 0;
  return Tap{base, to_u8(static_cast<int>(s0 * 3 + k - base)),
             to_u8(static_cast<int>(s1 * 3 + k - base)), frac, p.half_read,
             false};
}

// Column entry e (e < dst_width * channels) and row entry e (e < dst_height)
// of the generic path's tables.
__device__ __forceinline__ void resize_tables_elem(const GenericPlan &p, int2 *cols,
                                                   int2 *rows, int64_t e) {
  const int64_t n = static_cast<int64_t>(p.dst_width) * p.channels;
  if (e < n) {
    Tap t{};
    bool found = false;
    for (int64_t h = p.half_n - 1; h >= 0 && !found; --h) {
      const int64_t start = min(p.two_x * 32 + 8 * h, n - 8);
      if (e >= start && e < start + 8) {
        t = p.channels == 3 ? half_3(p, start, static_cast<int>(e - start))
                            : half_124(p, h, static_cast<int>(e - start));
        found = true;
      }
    }
    if (!found) {
      t = p.channels == 3 ? full_3(p, e / 16, static_cast<int>(e % 16))
                          : full_124(p, e / 16, static_cast<int>(e % 16));
    }
    // Taps stay inside the row even where the vector code would read beyond
    // it (only the kSetRightmostLanes lane reads outside its tbl range).
    const int64_t last = static_cast<int64_t>(p.src_width) * p.channels - 1;
    const int left = static_cast<int>(min(t.base + t.idx0, last));
    const int right = static_cast<int>(min(t.base + t.idx1, last));
    const int zero = (t.idx0 < t.load ? 0 : 1) |
                     ((t.idx1 < t.load || t.direct_right) ? 0 : 2);
    cols[e] = make_int2(left, ((right - left) << 16) | (zero << 8) | t.frac);
  }
  if (e < p.dst_height) {
    // ResizeGenericU8Operation::process_row (:730-754)
    const int64_t sy_fixp = aligned_scale(e, p.src_height, p.dst_height);
    const int64_t sy = sy_fixp >> 16;
    const int64_t max_sy = p.src_height - 1;
    const int top = static_cast<int>(min(max(sy, int64_t{0}), max_sy));
    const int bot = static_cast<int>(min(max(sy + 1, int64_t{0}), max_sy));
    const int yfrac = static_cast<int>((sy_fixp - sy * 65536) >> 8);
    rows[e] = make_int2(top, static_cast<int>(static_cast<unsigned>(bot) |
                                              (static_cast<unsigned>(yfrac) << 24)));
  }
}

// ---------------------------------------------------------------------------
// Exact 1-channel 2x / 4x upsizers: separable weights (3,1)/4 resp.
// (7,1),(5,3),(3,5),(1,7)/8 per axis with clamped taps and one rounding
// (the edge rows' 1-D formulas of resize_linear_neon.cpp are this closed form
// with the other axis' weight all on one tap).  Their tables use the same
// layout as the generic path's, with the right/bottom tap weight in place of
// the 8-bit fraction.
template <int F>
__device__ __forceinline__ void exact_taps(int d, int n, int &i0, int &i1,
                                           int &w1) {
  const int q = d - F / 2;
  if (q < 0) {
    i0 = i1 = 0;
    w1 = 0;
    return;
  }
  const int s = q / F, ph = q - s * F;
  i0 = s;
  i1 = min(s + 1, n - 1);
  w1 = F == 2 ? (ph == 0 ? 1 : 3) : 1 + 2 * ph;
}

template <int F>
__device__ __forceinline__ void exact_tables_elem(const ResizeArgs &a, int2 *cols,
                                                  int2 *rows, int e) {
  int i0, i1, w1;
  if (e < a.dst_width) {
    exact_taps<F>(e, a.src_width, i0, i1, w1);
    cols[e] = make_int2(i0, ((i1 - i0) << 16) | w1);
  }
  if (e < a.dst_height) {
    exact_taps<F>(e, a.src_height, i0, i1, w1);
    rows[e] = make_int2(i0, static_cast<int>(static_cast<unsigned>(i1) |
                                             (static_cast<unsigned>(w1) << 24)));
  }
}

// ---------------------------------------------------------------------------
// Staged gather.  A workgroup owns kTileE destination bytes of `band` rows.
// It reads its column taps once, finds the source byte span they cover and
// the source rows the band needs, copies that window into LDS with dword
// loads, then every output is four LDS byte reads and two (or one) lerps.
constexpr int kTileE = 4 * kThreads;       // destination bytes per workgroup
constexpr int kStageBytes = 24 * 1024;     // LDS window capacity

enum { kLerpGeneric = 0, kExact2 = 2, kExact4 = 4 };

// vraddhn_u16(a << 8, (b - a) * w): exact in 16 bits for 0 <= w <= 255.
__device__ __forceinline__ int lerp8(int a, int b, int w) {
  return (256 * a + (b - a) * w + 128) >> 8;
}

template <int MODE>
__device__ __forceinline__ uint32_t combine(int pa, int pb, int pc, int pd,
                                            int wx, int wy) {
  if (MODE == kLerpGeneric) {
    return static_cast<uint32_t>(lerp8(lerp8(pa, pc, wy), lerp8(pb, pd, wy), wx));
  } else {
    constexpr int kUnit = MODE == kExact2 ? 4 : 8;
    constexpr int kShift = MODE == kExact2 ? 4 : 6;
    const int acc = (kUnit - wy) * ((kUnit - wx) * pa + wx * pb) +
                    wy * ((kUnit - wx) * pc + wx * pd);
    return static_cast<uint32_t>((acc + (1 << (kShift - 1))) >> kShift);
  }
}

typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ us2 as_us2(uint32_t v) { return __builtin_bit_cast(us2, v); }
__device__ __forceinline__ uint32_t as_u32(us2 v) { return __builtin_bit_cast(uint32_t, v); }

// combine<MODE> on two taps at once in 16-bit halves (v_pk_* arithmetic).
// Every intermediate fits 16 bits: the lerp's 256a + (b-a)w + 128 lies in
// [0, 65408] and wraps exactly as the reference's vraddhn_u16 operands do;
// the exact upsizers' sums stay below 2^14.
template <int MODE>
__device__ __forceinline__ us2 combine2(us2 pa, us2 pb, us2 pc, us2 pd, us2 wx, us2 wy) {
  const us2 k256 = us2{256, 256}, k128 = us2{128, 128};
  if (MODE == kLerpGeneric) {
    // 256a + (b - a)w + 128 == a(256 - w) + bw + 128: two multiply-adds
    const us2 iwy = k256 - wy, iwx = k256 - wx;
    const us2 left = (pa * iwy + pc * wy + k128) >> 8;
    const us2 right = (pb * iwy + pd * wy + k128) >> 8;
    return (left * iwx + right * wx + k128) >> 8;
  } else {
    constexpr uint16_t kUnit = MODE == kExact2 ? 4 : 8;
    constexpr uint16_t kShift = MODE == kExact2 ? 4 : 6;
    const us2 unit = us2{kUnit, kUnit};
    const us2 h0 = pa * (unit - wx) + pb * wx;
    const us2 h1 = pc * (unit - wx) + pd * wx;
    const us2 rnd = us2{1 << (kShift - 1), 1 << (kShift - 1)};
    return (h0 * (unit - wy) + h1 * wy + rnd) >> kShift;
  }
}

__device__ __forceinline__ int block_reduce(int v, bool is_max, int *scratch) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off);
    v = is_max ? max(v, o) : min(v, o);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) scratch[wave] = v;
  __syncthreads();
  int r = scratch[0];
#pragma unroll
  for (int i = 1; i < kThreads / 64; ++i)
    r = is_max ? max(r, scratch[i]) : min(r, scratch[i]);
  __syncthreads();
  return r;
}

// Column taps of destination bytes e0 .. e0 + 3 from the generic table:
// left and right source bytes, horizontal weight, and flags (bit 0: the left
// tap reads 0, bit 1: the right tap reads 0; both set past the row's end).
struct QuadTaps {
  int left[4], right[4], wx[4], zero[4];
};

__device__ __forceinline__ QuadTaps load_taps(const int2 *__restrict__ cols, int e0, int n) {
  QuadTaps q;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int e = e0 + k;
    const int2 c = cols[min(e, n - 1)];
    q.left[k] = c.x;
    q.right[k] = c.x + (c.y >> 16);
    q.wx[k] = c.y & 0xFF;
    q.zero[k] = e < n ? (c.y >> 8) & 3 : 3;
  }
  return q;
}

// Least and greatest byte index read by the taps (tmax < 0: none read).
__device__ __forceinline__ void tap_span(const QuadTaps &q, int &tmin, int &tmax) {
  tmin = 0x7FFFFFFF;
  tmax = -1;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!(q.zero[k] & 1)) { tmin = min(tmin, q.left[k]); tmax = max(tmax, q.left[k]); }
    if (!(q.zero[k] & 2)) { tmin = min(tmin, q.right[k]); tmax = max(tmax, q.right[k]); }
  }
}

// Byte selector of v_perm_b32 over an 8-byte window {hi, lo} that starts at
// byte `base`: taps k0 and k0 + 1 of idx land in the low bytes of a 16-bit
// pair; 0x0C selects zero, which also realises the vector code's "index
// beyond the load reads 0".
__device__ __forceinline__ uint32_t pair_sel(const int *idx, const int *zero, int flag,
                                             int k0, int base) {
  uint32_t v = 0x0C000C00u;
  v |= (zero[k0] & flag) ? 0x0Cu : static_cast<uint32_t>(idx[k0] - base);
  v |= ((zero[k0 + 1] & flag) ? 0x0Cu : static_cast<uint32_t>(idx[k0 + 1] - base)) << 16;
  return v;
}

// The same over a 16-byte window: the selector for its low 8 bytes (upper
// false) or its high 8 bytes; taps in the other half select zero, so the
// two v_perm results are OR-ed.
__device__ __forceinline__ uint32_t pair_sel_half(const int *idx, const int *zero, int flag,
                                                  int k0, int base, bool upper) {
  uint32_t v = 0x0C000C00u;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = k0 + h;
    const int o = (zero[k] & flag) ? -1 : idx[k] - base;
    const bool mine = o >= 0 && (upper ? o >= 8 : o < 8);
    v |= static_cast<uint32_t>(mine ? (upper ? o - 8 : o) : 0x0C) << (16 * h);
  }
  return v;
}

// Four outputs from 8-byte windows {t1, t0} (top row) and {b1, b0} (bottom
// row); sel = {l01, l23, r01, r23}.
template <int MODE>
__device__ __forceinline__ uint32_t lerp_quad8(uint32_t t0, uint32_t t1, uint32_t b0,
                                               uint32_t b1, const uint32_t *sel, us2 wx01,
                                               us2 wx23, int wy) {
  const us2 wy2 = us2{static_cast<uint16_t>(wy), static_cast<uint16_t>(wy)};
  const us2 r01 = combine2<MODE>(as_us2(__builtin_amdgcn_perm(t1, t0, sel[0])),
                                 as_us2(__builtin_amdgcn_perm(t1, t0, sel[2])),
                                 as_us2(__builtin_amdgcn_perm(b1, b0, sel[0])),
                                 as_us2(__builtin_amdgcn_perm(b1, b0, sel[2])), wx01, wy2);
  const us2 r23 = combine2<MODE>(as_us2(__builtin_amdgcn_perm(t1, t0, sel[1])),
                                 as_us2(__builtin_amdgcn_perm(t1, t0, sel[3])),
                                 as_us2(__builtin_amdgcn_perm(b1, b0, sel[1])),
                                 as_us2(__builtin_amdgcn_perm(b1, b0, sel[3])), wx23, wy2);
  // bytes 0 and 2 of each pair word -> one dword
  return __builtin_amdgcn_perm(as_u32(r23), as_u32(r01), 0x06040200u);
}

// Four outputs from 16-byte windows t[0..3] and b[0..3]; wsel[s][half] for
// s = l01, l23, r01, r23.
template <int MODE>
__device__ __forceinline__ uint32_t lerp_quad16(const uint32_t *t, const uint32_t *b,
                                                const uint32_t (*wsel)[2], us2 wx01, us2 wx23,
                                                int wy) {
  auto pick = [&](const uint32_t *x, int s_) {
    return as_us2(__builtin_amdgcn_perm(x[1], x[0], wsel[s_][0]) |
                  __builtin_amdgcn_perm(x[3], x[2], wsel[s_][1]));
  };
  const us2 wy2 = us2{static_cast<uint16_t>(wy), static_cast<uint16_t>(wy)};
  const us2 r01 = combine2<MODE>(pick(t, 0), pick(t, 2), pick(b, 0), pick(b, 2), wx01, wy2);
  const us2 r23 = combine2<MODE>(pick(t, 1), pick(t, 3), pick(b, 1), pick(b, 3), wx23, wy2);
  return __builtin_amdgcn_perm(as_u32(r23), as_u32(r01), 0x06040200u);
}

// Destination bytes e0 .. e0 + 3 of row y (fewer at the row's end).
__device__ __forceinline__ void store_dst_quad(const ResizeArgs &a, int y, int e0, int n,
                                               uint32_t packed) {
  uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride;
  if (e0 + 3 < n && (reinterpret_cast<uintptr_t>(out + e0) & 3) == 0) {
    store_nt(reinterpret_cast<uint32_t *>(out + e0), packed);
  } else {
    for (int k = 0; k < 4 && e0 + k < n; ++k)
      out[e0 + k] = static_cast<uint8_t>(packed >> (8 * k));
  }
}

// Whether the generic lerp of `a` reads its taps straight from global
// memory (resize_direct_rows) instead of through an LDS stage: downsizing
// rows by 1.5 or more.
__host__ __device__ inline bool generic_reads_direct(const ResizeArgs &a) {
  return 2 * static_cast<int64_t>(a.src_height) >= 3 * static_cast<int64_t>(a.dst_height);
}

// Rows a thread of the direct path has in flight at once.
constexpr int kDirectRows = 4;

// Threads per destination row of a batched direct-path item: a power of two
// no larger than needed for the row's 4-byte quads; the workgroup's other
// threads take further rows of the band.
__host__ __device__ inline int direct_lanes(const ResizeArgs &a) {
  const int quads = (a.dst_width * a.channels + 3) / 4;
  int lanes = kThreads;
  while (lanes > 1 && lanes / 2 >= quads) lanes /= 2;
  return lanes;
}

// Band of a direct-path work item: one group of kDirectRows rows for each
// thread (rows_per_pass = workgroup threads / threads per row).
inline int direct_band_rows(int rows_per_pass = 1) { return kDirectRows * rows_per_pass; }

// Generic lerp read straight from global memory: destination bytes
// e0 .. e0 + 3 of rows y_begin, y_begin + y_step, ... < y_end.  Used for
// downsizing by 1.5 or more, where a source row serves about one destination
// row and an LDS stage would buy no reuse: each row is one 8- or 16-byte
// load at any alignment (gfx950 vector memory runs in unaligned mode), the
// window start pulled back from the row's end so that it stays inside the
// row; taps spread wider than 16 bytes read byte by byte.  The loads of
// kDirectRows rows are issued together (rows past the end repeat the last
// row and are not stored), so a band costs one memory round trip, not one
// per row.
template <int MODE>
__device__ __forceinline__ void resize_direct_rows(const ResizeArgs &a,
                                                   const int2 *__restrict__ cols,
                                                   const int2 *__restrict__ rows, int e0,
                                                   int y_begin, int y_end, int y_step) {
  const int n = a.dst_width * a.channels;
  if (e0 >= n || y_begin >= y_end) return;
  const int row_bytes = a.src_width * a.channels;
  const QuadTaps q = load_taps(cols, e0, n);
  int tmin, tmax;
  tap_span(q, tmin, tmax);
  const bool narrow = tmax >= 0 && tmax - tmin < 8 && row_bytes >= 8;
  const bool wide = tmax >= 0 && !narrow && tmax - tmin < 16 && row_bytes >= 16;
  // taps lie in [tmin, tmax] within [0, row_bytes): offsets from `base`
  // stay below the window size
  const int base = narrow ? min(tmin, row_bytes - 8) : wide ? min(tmin, row_bytes - 16) : 0;
  uint32_t sel[4], wsel[4][2];
  if (narrow) {
    sel[0] = pair_sel(q.left, q.zero, 1, 0, base);
    sel[1] = pair_sel(q.left, q.zero, 1, 2, base);
    sel[2] = pair_sel(q.right, q.zero, 2, 0, base);
    sel[3] = pair_sel(q.right, q.zero, 2, 2, base);
  } else if (wide) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      wsel[0][u] = pair_sel_half(q.left, q.zero, 1, 0, base, u);
      wsel[1][u] = pair_sel_half(q.left, q.zero, 1, 2, base, u);
      wsel[2][u] = pair_sel_half(q.right, q.zero, 2, 0, base, u);
      wsel[3][u] = pair_sel_half(q.right, q.zero, 2, 2, base, u);
    }
  }
  const us2 wx01 = us2{static_cast<uint16_t>(q.wx[0]), static_cast<uint16_t>(q.wx[1])};
  const us2 wx23 = us2{static_cast<uint16_t>(q.wx[2]), static_cast<uint16_t>(q.wx[3])};
  auto row_of = [&](int y, int &wy) {
    const int2 r = rows[min(y, y_end - 1)];
    wy = (r.y >> 24) & 0xFF;
    return r;
  };
  auto row_ptr = [&](int i) { return a.src + static_cast<size_t>(i) * a.src_stride; };
  if (narrow || wide) {
    for (int y = y_begin; y < y_end; y += kDirectRows * y_step) {
      int wy[kDirectRows];
      uint4 tv[kDirectRows], bv[kDirectRows];
#pragma unroll
      for (int i = 0; i < kDirectRows; ++i) {
        const int2 r = row_of(y + i * y_step, wy[i]);
        const uint8_t *tp = row_ptr(r.x) + base, *bp = row_ptr(r.y & 0xFFFFFF) + base;
        if (narrow) {
          const uint2 t = *reinterpret_cast<const u2_any *>(tp);
          const uint2 b = *reinterpret_cast<const u2_any *>(bp);
          tv[i] = make_uint4(t.x, t.y, 0, 0);
          bv[i] = make_uint4(b.x, b.y, 0, 0);
        } else {
          tv[i] = *reinterpret_cast<const u4_any *>(tp);
          bv[i] = *reinterpret_cast<const u4_any *>(bp);
        }
      }
#pragma unroll
      for (int i = 0; i < kDirectRows; ++i) {
        uint32_t packed;
        if (narrow) {
          packed = lerp_quad8<MODE>(tv[i].x, tv[i].y, bv[i].x, bv[i].y, sel, wx01, wx23, wy[i]);
        } else {
          const uint32_t t[4] = {tv[i].x, tv[i].y, tv[i].z, tv[i].w};
          const uint32_t b[4] = {bv[i].x, bv[i].y, bv[i].z, bv[i].w};
          packed = lerp_quad16<MODE>(t, b, wsel, wx01, wx23, wy[i]);
        }
        if (y + i * y_step < y_end) store_dst_quad(a, y + i * y_step, e0, n, packed);
      }
    }
    return;
  }
  for (int y = y_begin; y < y_end; y += y_step) {
    int wy;
    const int2 r = row_of(y, wy);
    const uint8_t *tp = row_ptr(r.x), *bp = row_ptr(r.y & 0xFFFFFF);
    uint32_t packed = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int pa = (q.zero[k] & 1) ? 0 : tp[q.left[k]];
      const int pc = (q.zero[k] & 1) ? 0 : bp[q.left[k]];
      const int pb = (q.zero[k] & 2) ? 0 : tp[q.right[k]];
      const int pd = (q.zero[k] & 2) ? 0 : bp[q.right[k]];
      packed |= (combine<MODE>(pa, pb, pc, pd, q.wx[k], wy) & 0xFFu) << (8 * k);
    }
    store_dst_quad(a, y, e0, n, packed);
  }
}

// LDS of one staged tile: the window plus 16 spare bytes (the 8- and 16-byte
// reads of the narrow and wide paths may start in the window's last dwords)
// and the reduction scratch.
constexpr int kStagedSmem = kStageBytes + 16 + 4 * (kThreads / 64);

// Destination bytes [bx*kTileE, +kTileE) of rows [by*band, +band).
// `lanes` threads (a power of two) share each destination row: kThreads for
// rows of a whole tile or more; for narrower rows fewer, and the band's
// rows are dealt to the kThreads / lanes groups of them.
template <int MODE>
__device__ __forceinline__ void resize_staged_tile(const ResizeArgs &a,
                                                   const int2 *__restrict__ cols,
                                                   const int2 *__restrict__ rows,
                                                   int band, int bx, int by,
                                                   uint8_t *smem, int lanes = kThreads) {
  const int n = a.dst_width * a.channels;
  const int row_bytes = a.src_width * a.channels;
  const int e0 = bx * kTileE + static_cast<int>(threadIdx.x % lanes) * 4;
  const int row_lane = static_cast<int>(threadIdx.x / lanes), row_step = kThreads / lanes;
  const int y_begin = by * band;
  const int y_end = min(y_begin + band, a.dst_height);
  // The choice depends on the image only, so the whole workgroup takes it
  // and the direct path (no barriers, no LDS) never splits a barrier.
  if (MODE == kLerpGeneric && generic_reads_direct(a)) {
    resize_direct_rows<MODE>(a, cols, rows, e0, y_begin + row_lane, y_end, row_step);
    return;
  }
  uint8_t *win = smem;
  int *scratch = reinterpret_cast<int *>(smem + kStageBytes + 16);
  __syncthreads();  // LDS may still hold the previous tile of a batched loop

  QuadTaps q = load_taps(cols, e0, n);
  int lo, hi;
  tap_span(q, lo, hi);
  lo = block_reduce(lo, false, scratch);
  hi = block_reduce(hi, true, scratch);
  if (hi < lo) { lo = 0; hi = 0; }
  const int lo4 = lo & ~3;
  const int pitch = ((hi + 1 - lo4) + 15) & ~15;  // whole 16-byte chunks

  const int r_lo = rows[y_begin].x;
  const int r_hi = rows[y_end - 1].y & 0xFFFFFF;
  const int nrows = r_hi - r_lo + 1;
  const bool staged = nrows * pitch <= kStageBytes;

  if (staged) {
    // 16-byte loads at any alignment (unaligned mode); the row's last
    // partial chunk byte by byte
    const int chunks = pitch >> 4;
    for (int i = threadIdx.x; i < nrows * chunks; i += kThreads) {
      const int r = i / chunks, c = i - r * chunks;
      const uint8_t *srow = a.src + static_cast<size_t>(r_lo + r) * a.src_stride;
      const int off = lo4 + 16 * c;
      uint4 v;
      if (off + 16 <= row_bytes) {
        v = *reinterpret_cast<const u4_any *>(srow + off);
      } else {
        uint32_t w[4] = {0, 0, 0, 0};
        for (int b = 0; b < 16 && off + b < row_bytes; ++b)
          w[b >> 2] |= static_cast<uint32_t>(srow[off + b]) << (8 * (b & 3));
        v = make_uint4(w[0], w[1], w[2], w[3]);
      }
      *reinterpret_cast<uint4 *>(win + r * pitch + 16 * c) = v;
    }
    __syncthreads();
  }
  if (e0 >= n) return;  // no barrier follows in this tile
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    q.left[k] -= lo4;
    q.right[k] -= lo4;
  }

  // When a thread's taps lie within 8 bytes of a dword boundary (upsizing,
  // mild downsizing) each source row is one 8-byte LDS read plus bit-field
  // extracts instead of eight byte reads.
  int tmin, tmax;
  tap_span(q, tmin, tmax);
  const int tbase = tmin & ~3;
  const bool narrow = staged && tmax >= 0 && tmax - tbase < 8;
  // Taps within 16 bytes (2:1 downsizing of 1 and 3 channels): four dword
  // reads per row; each tap pair is picked from the low or the high 8 bytes
  // by two v_perm_b32 whose other selectors read zero, then OR-ed.
  const bool wide = staged && !narrow && tmax >= 0 && tmax - tbase < 16;
  uint32_t sel[4], wsel[4][2];  // l01, l23, r01, r23 (x {low, high 8 bytes})
  if (narrow) {
    sel[0] = pair_sel(q.left, q.zero, 1, 0, tbase);
    sel[1] = pair_sel(q.left, q.zero, 1, 2, tbase);
    sel[2] = pair_sel(q.right, q.zero, 2, 0, tbase);
    sel[3] = pair_sel(q.right, q.zero, 2, 2, tbase);
  } else if (wide) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      wsel[0][u] = pair_sel_half(q.left, q.zero, 1, 0, tbase, u);
      wsel[1][u] = pair_sel_half(q.left, q.zero, 1, 2, tbase, u);
      wsel[2][u] = pair_sel_half(q.right, q.zero, 2, 0, tbase, u);
      wsel[3][u] = pair_sel_half(q.right, q.zero, 2, 2, tbase, u);
    }
  }
  const us2 wx01 = us2{static_cast<uint16_t>(q.wx[0]), static_cast<uint16_t>(q.wx[1])};
  const us2 wx23 = us2{static_cast<uint16_t>(q.wx[2]), static_cast<uint16_t>(q.wx[3])};

  for (int y = y_begin + row_lane; y < y_end; y += row_step) {
    const int2 r = rows[y];
    const int top = r.x, bot = r.y & 0xFFFFFF;
    const int wy = (r.y >> 24) & 0xFF;
    uint32_t packed = 0;
    if (narrow || wide) {
      const uint32_t *tp = reinterpret_cast<const uint32_t *>(win + (top - r_lo) * pitch + tbase);
      const uint32_t *bp = reinterpret_cast<const uint32_t *>(win + (bot - r_lo) * pitch + tbase);
      if (narrow) {
        packed = lerp_quad8<MODE>(tp[0], tp[1], bp[0], bp[1], sel, wx01, wx23, wy);
      } else {
        const uint32_t t[4] = {tp[0], tp[1], tp[2], tp[3]};
        const uint32_t b[4] = {bp[0], bp[1], bp[2], bp[3]};
        packed = lerp_quad16<MODE>(t, b, wsel, wx01, wx23, wy);
      }
    } else {
      // per-byte reads from the LDS window, or from global memory where the
      // window does not fit
      const uint8_t *t = staged ? win + (top - r_lo) * pitch
                                : a.src + static_cast<size_t>(top) * a.src_stride + lo4;
      const uint8_t *b = staged ? win + (bot - r_lo) * pitch
                                : a.src + static_cast<size_t>(bot) * a.src_stride + lo4;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int pa = (q.zero[k] & 1) ? 0 : t[q.left[k]];
        const int pc = (q.zero[k] & 1) ? 0 : b[q.left[k]];
        const int pb = (q.zero[k] & 2) ? 0 : t[q.right[k]];
        const int pd = (q.zero[k] & 2) ? 0 : b[q.right[k]];
        packed |= (combine<MODE>(pa, pb, pc, pd, q.wx[k], wy) & 0xFFu) << (8 * k);
      }
    }
    store_dst_quad(a, y, e0, n, packed);
  }
}

// ---------------------------------------------------------------------------
// Exact 0.5 x 0.5, one kernel per channel count.  Destination bytes per thread: 16 for 1, 2 and 4 channels, 24 (eight
// pixels) for 3, so that a thread's source bytes are whole 16-byte chunks.
__host__ __device__ constexpr int quarter_thread_bytes(int channels) {
  return channels == 3 ? 24 : 16;
}

// Destination bytes [e0, e0 + OB) of rows y_begin, y_begin + y_step, ...
// with e0 = (bx * lanes + lane) * OB: `lanes` threads share a row (all of the
// workgroup for wide images; fewer, with the rest on other rows, for narrow
// ones).  Each output byte is (a + b + c + d + 2) >> 2 of its 2 x 2 source
// bytes.  Rows are read with 16-byte loads at any alignment (gfx950 vector
// memory is byte-addressed in unaligned mode); the last partial group of a
// row goes byte by byte.
template <int C>
__device__ __forceinline__ void resize_quarter_rows(const ResizeArgs &a, int bx, int lane,
                                                    int lanes, int y_begin, int y_end,
                                                    int y_step) {
  constexpr int OB = quarter_thread_bytes(C), IB = 2 * OB;
  typedef uint4 __attribute__((aligned(1))) u4u;
  const int n = a.dst_width * C;
  const int e0 = (bx * lanes + lane) * OB;
  if (e0 >= n) return;
  const bool whole = e0 + OB <= n;
  for (int y = y_begin; y < y_end; y += y_step) {
    const uint8_t *r0 = a.src + static_cast<size_t>(2 * y) * a.src_stride + 2 * e0;
    const uint8_t *r1 = r0 + a.src_stride;
    uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride + e0;
    if (whole) {
      uint32_t t[IB / 4], b[IB / 4];
#pragma unroll
      for (int k = 0; k < IB / 16; ++k) {
        const uint4 qt = *reinterpret_cast<const u4u *>(r0 + 16 * k);
        const uint4 qb = *reinterpret_cast<const u4u *>(r1 + 16 * k);
        t[4 * k] = qt.x; t[4 * k + 1] = qt.y; t[4 * k + 2] = qt.z; t[4 * k + 3] = qt.w;
        b[4 * k] = qb.x; b[4 * k + 1] = qb.y; b[4 * k + 2] = qb.z; b[4 * k + 3] = qb.w;
      }
      auto byte = [](const uint32_t *v, int i) { return (v[i >> 2] >> (8 * (i & 3))) & 0xFFu; };
      uint32_t o[OB / 4];
#pragma unroll
      for (int w = 0; w < OB / 4; ++w) {
        uint32_t packed = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int j = 4 * w + k;                  // output byte in the group
          const int i0 = 2 * (j - j % C) + j % C;   // its left source byte
          const uint32_t sum = byte(t, i0) + byte(t, i0 + C) + byte(b, i0) + byte(b, i0 + C);
          packed |= ((sum + 2) >> 2) << (8 * k);
        }
        o[w] = packed;
      }
      if ((reinterpret_cast<uintptr_t>(out) & 15) == 0 && OB == 16) {
        store_nt(reinterpret_cast<uint4 *>(out), make_uint4(o[0], o[1], o[2], o[3]));
      } else if ((reinterpret_cast<uintptr_t>(out) & 7) == 0) {
#pragma unroll
        for (int w = 0; w < OB / 8; ++w)
          store_nt(reinterpret_cast<uint2 *>(out) + w, make_uint2(o[2 * w], o[2 * w + 1]));
      } else {
        typedef uint32_t __attribute__((aligned(1))) u32u;
#pragma unroll
        for (int w = 0; w < OB / 4; ++w) reinterpret_cast<u32u *>(out)[w] = o[w];
      }
    } else {
      for (int k = 0; e0 + k < n; ++k) {
        const int i0 = 2 * (k - k % C) + k % C;
        const uint32_t sum = r0[i0] + r0[i0 + C] + r1[i0] + r1[i0 + C];
        out[k] = static_cast<uint8_t>((sum + 2) >> 2);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Generic path at exactly 2x in both axes, 4 channels.  Every coordinate of
// the generic generator is then exact: output 2m takes source m-1 and m with
// fraction 192, output 2m+1 takes m and m+1 with fraction 64, indices
// clamped to the image (the generic tables give the same clamped taps, and
// lerping a value with itself returns it).  L(a, b, 192) = (a + 3b + 2) >> 2
// and L(a, b, 64) = (3a + b + 2) >> 2 exactly.  Per byte, (3a + b + 2) >> 2
// equals (a + m + r) >> 1 with the rounded mean m = (a + b + 1) >> 1 and
// r = 1 when a + b is even (the tests check this over all byte pairs), and
// v_lerp_u8 computes (x + y + r) >> 1 on the four bytes of a dword at once.
// So a lerp of a 4-channel pixel is one v_lerp_u8 with the pair's mean,
// and the mean (one v_lerp_u8 and one v_bitop3_b32) serves both lerps of
// the pair, L(a, b, 64) and L(a, b, 192).
//
// Row group g in [0, h] produces output rows 2g-1 (vertical fraction 64) and
// 2g (fraction 192) from source rows max(g-1, 0) and min(g, h-1); group 0
// has only row 0 and group h only row 2h-1.  A thread owns 2 source pixels
// of its group, i.e. 4 output pixels (16 bytes) of each output row, so the
// lanes of a wave store one contiguous 1 KiB run per output row.
constexpr int kUp2x4ThreadPx = 2;

// Rounded mean of the bytes of a and b, and the rounding bits of the
// three-to-one lerps of the pair.
struct ByteMean {
  uint32_t m, r;
};

__device__ __forceinline__ ByteMean byte_mean(uint32_t a, uint32_t b) {
  return {__builtin_amdgcn_lerp(a, b, 0x01010101u), ~(a ^ b) & 0x01010101u};
}

// (3a + b + 2) >> 2 per byte, given the mean of a and b.
__device__ __forceinline__ uint32_t lerp_q3(uint32_t a, const ByteMean &ab) {
  return __builtin_amdgcn_lerp(a, ab.m, ab.r);
}

__device__ __forceinline__ void up2x4_group(const ResizeArgs &a, int bx, int g) {
  const int w = a.src_width, h = a.src_height;
  const int m0 = (bx * kThreads + threadIdx.x) * kUp2x4ThreadPx;
  if (m0 >= w || g > h) return;
  const uint8_t *top = a.src + static_cast<size_t>(g > 0 ? g - 1 : 0) * a.src_stride;
  const uint8_t *bot = a.src + static_cast<size_t>(g < h ? g : h - 1) * a.src_stride;
  const bool al = ((reinterpret_cast<uintptr_t>(a.src) | a.src_stride) & 7) == 0;
  // source pixels m0-1 .. m0+2 (clamped) of both rows
  uint32_t t[4], b[4];
  auto px = [&](const uint8_t *row, int m) -> uint32_t {
    m = min(max(m, 0), w - 1);
    const uint8_t *p = row + 4 * m;
    if ((reinterpret_cast<uintptr_t>(p) & 3) == 0) return *reinterpret_cast<const uint32_t *>(p);
    return static_cast<uint32_t>(p[0]) | (static_cast<uint32_t>(p[1]) << 8) |
           (static_cast<uint32_t>(p[2]) << 16) | (static_cast<uint32_t>(p[3]) << 24);
  };
  if (al && m0 + 2 <= w) {
    const uint2 qt = *reinterpret_cast<const uint2 *>(top + 4 * m0);
    const uint2 qb = *reinterpret_cast<const uint2 *>(bot + 4 * m0);
    t[1] = qt.x; t[2] = qt.y;
    b[1] = qb.x; b[2] = qb.y;
  } else {
    t[1] = px(top, m0); t[2] = px(top, m0 + 1);
    b[1] = px(bot, m0); b[2] = px(bot, m0 + 1);
  }
  // Pixels m0 - 1 and m0 + 2 are the neighbouring lanes' pixels (DPP wave
  // shifts); only the wave's end lanes and the image edges load them.
  // (A neighbour lane that exited past the image end is never read: the
  // lane before it is then at the edge and loads.)
  const int lane = static_cast<int>(threadIdx.x & 63);
  const uint32_t tl = __builtin_amdgcn_update_dpp(0, t[2], 0x138, 0xF, 0xF, false);  // wave_shr:1
  const uint32_t bl = __builtin_amdgcn_update_dpp(0, b[2], 0x138, 0xF, 0xF, false);
  const uint32_t tr = __builtin_amdgcn_update_dpp(0, t[1], 0x130, 0xF, 0xF, false);  // wave_shl:1
  const uint32_t br = __builtin_amdgcn_update_dpp(0, b[1], 0x130, 0xF, 0xF, false);
  if (lane == 0 || m0 == 0) {
    t[0] = px(top, m0 - 1);
    b[0] = px(bot, m0 - 1);
  } else {
    t[0] = tl;
    b[0] = bl;
  }
  if (lane == 63 || m0 + 2 >= w) {
    t[3] = px(top, m0 + 2);
    b[3] = px(bot, m0 + 2);
  } else {
    t[3] = tr;
    b[3] = br;
  }
  const int n_out = min(4, 2 * (w - m0));  // output pixels of this thread per row
  const bool out_al = ((reinterpret_cast<uintptr_t>(a.dst) | a.dst_stride) & 15) == 0;
  // vertical lerps: v64 = L(top, bottom, 64) for row 2g-1, v192 for row 2g
  uint32_t v64[4], v192[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const ByteMean tb = byte_mean(t[k], b[k]);
    v64[k] = lerp_q3(t[k], tb);
    v192[k] = lerp_q3(b[k], tb);
  }
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int y = 2 * g - 1 + half;  // half 0: fraction 64, half 1: fraction 192
    if (y < 0 || y >= a.dst_height) continue;
    const uint32_t *v = half == 0 ? v64 : v192;
    // output 2m: L(V[m-1], V[m], 192); output 2m+1: L(V[m], V[m+1], 64)
    const ByteMean p01 = byte_mean(v[0], v[1]), p12 = byte_mean(v[1], v[2]),
                   p23 = byte_mean(v[2], v[3]);
    const uint32_t o[4] = {lerp_q3(v[1], p01), lerp_q3(v[1], p12), lerp_q3(v[2], p12),
                           lerp_q3(v[2], p23)};
    uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride + 8 * static_cast<size_t>(m0);
    if (n_out == 4 && out_al) {
      store_nt(reinterpret_cast<uint4 *>(out), make_uint4(o[0], o[1], o[2], o[3]));
    } else {
      for (int k = 0; k < n_out; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) out[4 * k + c] = static_cast<uint8_t>(o[k] >> (8 * c));
    }
  }
}

enum class Path { kNone, kQuarter, kUp2, kUp4, kUp2x4, kGeneric };

inline Path pick(const ResizeArgs &a) {
  const int64_t sw = a.src_width, sh = a.src_height, dw = a.dst_width,
                dh = a.dst_height;
  if (sw == 0 || sh == 0) return Path::kNone;
  if (sw == 2 * dw && sh == 2 * dh) return Path::kQuarter;
  if (a.channels == 1 && sw * 2 == dw && sh * 2 == dh) return Path::kUp2;
  if (a.channels == 1 && sw * 4 == dw && sh * 4 == dh) return Path::kUp4;
  if (a.channels == 4 && sw * 2 == dw && sh * 2 == dh) return Path::kUp2x4;
  return Path::kGeneric;
}

// Rows per workgroup band so that the staged window fits in LDS for a
// typical tile; tiles whose window does not fit read from global memory.
inline int band_rows(const ResizeArgs &a) {
  const double sx = static_cast<double>(a.src_width) / a.dst_width;
  const double sy = static_cast<double>(a.src_height) / a.dst_height;
  const double pitch = kTileE * sx + 8.0 * a.channels + 8.0;
  const double rows_fit = kStageBytes / pitch;
  int band = static_cast<int>((rows_fit - 2.0) / (sy > 0 ? sy : 1.0)) + 1;
  band = band < 1 ? 1 : band > 64 ? 64 : band;
  return band;
}

}  // namespace rsz
}  // namespace kcv

#endif  // KCV_RESIZE_BODY_H
