// Device bodies of the tiled 2-D filters (separable integer filters and the
// median), shared by the per-image kernels (filters.hip) and the batched
// persistent kernels (batch.hip).
//
// One workgroup of 256 threads (4 wave64) produces a tile of 256 bytes x 16
// rows of the interleaved u8 image.  The tile's source window, widened by the
// filter radius in rows and by a 16-byte-aligned pad in bytes, is staged in
// LDS once; the border mode is applied while staging, so the arithmetic that
// follows never tests for borders.  Interior windows are fetched with
// 16-byte loads, edge windows byte by byte through the border map.
//
// Reference semantics:
//   borders      include/kleidicv/workspace/border_{3x3,5x5,7x7}.h,
//                src/filters/median_blur_border_handling.h:14-56
//   binomial     src/filters/gaussian_blur_fixed_neon.cpp:27-459
//   sigma        src/filters/gaussian_blur_fixed_neon.cpp:461-529
//   separable    src/filters/separable_filter_2d_neon.cpp:19-119
//   median       src/filters/median_blur_sorting_network_neon.cpp:173-204
#ifndef KCV_FILTERS_BODY_H
#define KCV_FILTERS_BODY_H

#include "kcv_internal.h"

namespace kcv {
namespace filt {


constexpr int kTileW = 256;
constexpr int kTileH = 16;
constexpr int kThreads = 256;

// Border index map with a single reflection (valid for overhang <= n, the
// reference's w,h >= k-1 precondition).  The final clamp only affects window
// positions no output reads (tiles hanging over the image's right/bottom).
__device__ __forceinline__ int map_index(int i, int n, int border) {
  if (static_cast<unsigned>(i) < static_cast<unsigned>(n)) return i;
  int r;
  switch (border) {
    case kReflect:
      r = i < 0 ? -i - 1 : 2 * n - i - 1;
      break;
    case kWrap:
      r = i < 0 ? n + i : i - n;
      break;
    case kReverse:
      r = i < 0 ? min(-i, n - 1) : 2 * n - i - 2;
      break;
    default:
      r = i < 0 ? 0 : n - 1;
      break;
  }
  return min(max(r, 0), n - 1);
}

// LDS window of a tile of TH output rows for a radius-R filter.
template <int R, int TH = kTileH>
struct Window {
  static constexpr int kPad = 16 * ((R + 3) / 4);  // >= R*C bytes for C <= 4
  static constexpr int kCols = kTileW + 2 * kPad;
  static constexpr int kRows = TH + 2 * R;
};

// Column border map as one branch-free formula for the four filter
// borders: a position left of the row maps to sl * x + bl, one right of it
// to sr * x + br, then everything is clamped (REPLICATE is the clamp alone).
// Equal to map_index for overhangs up to the row length.
struct ColMap {
  int sl, bl, sr, br, n;
};

__device__ __forceinline__ ColMap col_map(int border, int n) {
  switch (border) {
    case kReflect: return ColMap{-1, -1, -1, 2 * n - 1, n};
    case kWrap: return ColMap{1, n, 1, -n, n};
    case kReverse: return ColMap{-1, 0, -1, 2 * n - 2, n};
    default: return ColMap{0, 0, 0, n - 1, n};
  }
}

__device__ __forceinline__ int map_col(const ColMap &m, int x) {
  const int r = x < 0 ? m.sl * x + m.bl : (x >= m.n ? m.sr * x + m.br : x);
  return min(max(r, 0), m.n - 1);
}

// Stages rows [y0-R, y0+TH+R) x bytes [e0-pad, e0+kTileW+pad) of the
// border-extended image into `win` (row pitch Window<R>::kCols).  Rows are
// mapped through the border first, so every 16-byte chunk that lies inside
// the row is one 16-byte load (edge tiles and unaligned rows included); only
// chunks that cross the left or right edge are assembled byte by byte,
// through the branch-free column map.  The two kinds go to separate
// registers: all of a thread's 16-byte loads are issued first and stay in
// flight while the edge bytes are fetched, and nothing waits on a load
// before the LDS stores.
template <int R, int C, int TH = kTileH>
__device__ __forceinline__ void stage_window(uint8_t *win, const uint8_t *src,
                                             size_t stride, int width,
                                             int height, int border, int e0, int y0) {
  using G = Window<R, TH>;
  constexpr int kChunks = G::kCols / 16;
  constexpr int kTotal = G::kRows * kChunks;
  constexpr int kPerThread = (kTotal + kThreads - 1) / kThreads;
  // rows and strides of any alignment: 16-byte loads at byte addresses
  // (gfx950 vector memory runs in unaligned mode)
  typedef uint4 __attribute__((aligned(1))) u4u;
  const int row_bytes = width * C;
  const int c0 = e0 - G::kPad;
  uint4 v[kPerThread], ev[kPerThread];
  bool inside[kPerThread];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int i = threadIdx.x + k * kThreads;
    const int r = i / kChunks, q = i - r * kChunks;
    const uint8_t *srow = src + static_cast<size_t>(map_index(y0 - R + r, height, border)) * stride;
    const int e = c0 + 16 * q;
    inside[k] = i < kTotal && e >= 0 && e + 16 <= row_bytes;
    v[k] = inside[k] ? *reinterpret_cast<const u4u *>(srow + e) : make_uint4(0, 0, 0, 0);
  }
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int i = threadIdx.x + k * kThreads;
    const int r = i / kChunks, q = i - r * kChunks;
    const int e = c0 + 16 * q;
    ev[k] = make_uint4(0, 0, 0, 0);
    if (i < kTotal && !inside[k]) {
      const uint8_t *srow = src + static_cast<size_t>(map_index(y0 - R + r, height, border)) *
                                      stride;
      const ColMap cm = col_map(border, width);
      uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        const int eb = e + b;
        const int x = eb >= 0 ? eb / C : -((C - 1 - eb) / C);
        const int c = eb - x * C;
        w[b >> 2] |= static_cast<uint32_t>(srow[map_col(cm, x) * C + c]) << (8 * (b & 3));
      }
      ev[k] = make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int i = threadIdx.x + k * kThreads;
    if (i < kTotal) {
      const int r = i / kChunks, q = i - r * kChunks;
      // a bitwise blend rather than a select: the two arrays stay in registers
      const uint32_t m = inside[k] ? 0xFFFFFFFFu : 0u;
      *reinterpret_cast<uint4 *>(win + r * G::kCols + 16 * q) =
          make_uint4((v[k].x & m) | ev[k].x, (v[k].y & m) | ev[k].y, (v[k].z & m) | ev[k].z,
                     (v[k].w & m) | ev[k].w);
    }
  }
}

__device__ __forceinline__ void store_quad(uint8_t *row, int e, int row_bytes,
                                           uint32_t packed) {
  typedef uint32_t __attribute__((aligned(1))) u32u;
  if (e + 3 < row_bytes && (reinterpret_cast<uintptr_t>(row + e) & 3) == 0) {
    store_nt(reinterpret_cast<uint32_t *>(row + e), packed);
  } else if (e + 3 < row_bytes) {
    *reinterpret_cast<u32u *>(row + e) = packed;  // any alignment (unaligned mode)
  } else {
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (e + b < row_bytes) row[e + b] = static_cast<uint8_t>(packed >> (8 * b));
  }
}

// ---------------------------------------------------------------------------
// Separable integer filters.
//
// Vertical pass: each work item takes one dword (4 bytes) of K window rows
// and accumulates in 16-bit halves (bytes 0/2 and 1/3 side by side), then
// writes the four column sums as u16 to LDS with one 8-byte store.  Every
// vertical sum fits 16 bits: binomial and sigma taps sum to at most 256 and
// the saturating mode clamps at 65535, which leaves min(255, ...) unchanged
// (a clamped column only occurs where some row product already exceeds 255,
// and a zero horizontal tap ignores it either way).
//
// Horizontal pass: a thread produces 8 consecutive output bytes of one row.
// It reads the span of column sums those outputs need with 16-byte LDS
// loads (neighbouring lanes read neighbouring chunks: no bank conflicts) and
// accumulates in 32 bits with compile-time tap offsets (channel count is a
// template parameter).  Taps are at most 16 bits and column sums 16 bits:
// adjacent taps go two at a time through v_dot2_u32_u16 (their column sums
// packed into one dword, by v_perm_b32 when more than one channel puts them
// apart), and the last odd tap through the full-rate 24-bit multiplier.
typedef unsigned short sep_us2 __attribute__((ext_vector_type(2)));

// TH: output rows of a tile (sep_tile_rows(K) for single-image launches,
// kTileH in batches, where many images already fill the GPU and taller
// tiles load fewer halo rows).
template <int K, int TH>
using SepWindow = Window<K / 2, TH>;

template <int K, int TH>
constexpr int separable_win_bytes() {
  return (SepWindow<K, TH>::kRows * SepWindow<K, TH>::kCols + 15) / 16 * 16;
}

template <int K, int TH>
constexpr int separable_smem_bytes() {
  return separable_win_bytes<K, TH>() + 2 * TH * SepWindow<K, TH>::kCols;
}

// Whether the horizontal pass can accumulate in 16-bit lanes: binomial 3x3
// and 5x5 sums stay below 2^16 (at most 255 * 16 * 16 + 128 for 5x5), and
// the saturating mode's clamp at 65535 is left unchanged by its final
// min(255, ...) (every term is non-negative).  Binomial 7x7 and 9x9 and the
// sigma taps (whose error-diffused sum can reach 257) keep 32-bit sums.
template <int MODE, int K>
constexpr bool sep_horizontal_u16() {
  return MODE == kSepSaturate || (MODE == kSepBinomial && K <= 5);
}

template <int MODE, int K, int C, int TH>
__device__ __forceinline__ void separable_horizontal(const SepArgs &a, const uint16_t *vert,
                                                     int e0, int y0) {
  constexpr int R = K / 2;
  using G = SepWindow<K, TH>;
  constexpr int RC = R * C;
  constexpr int D = (RC + 7) / 8;         // 8-value chunks each side
  constexpr int NV = 8 * (2 * D + 1);      // column sums held per thread
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int row_bytes = a.width * C;
  uint32_t w[K];
#pragma unroll
  for (int j = 0; j < K; ++j) w[j] = static_cast<uint32_t>(a.wx[j]);
  sep_us2 wpair[K / 2];  // taps (2i, 2i + 1) for the dot2 form
#pragma unroll
  for (int i = 0; i < K / 2; ++i)
    wpair[i] = sep_us2{static_cast<uint16_t>(w[2 * i]), static_cast<uint16_t>(w[2 * i + 1])};
#pragma unroll
  for (int pass = 0; pass < TH / 8; ++pass) {
    const int r = ty + 8 * pass;
    const int y = y0 + r;
    if (y >= a.height) break;
    const uint4 *row = reinterpret_cast<const uint4 *>(vert + r * G::kCols);
    uint32_t v[NV / 2];
#pragma unroll
    for (int d = 0; d < 2 * D + 1; ++d) {
      const uint4 q = row[tx + G::kPad / 8 + d - D];
      v[4 * d] = q.x;
      v[4 * d + 1] = q.y;
      v[4 * d + 2] = q.z;
      v[4 * d + 3] = q.w;
    }
    auto col = [&](int i) -> uint32_t { return (v[i >> 1] >> (16 * (i & 1))) & 0xFFFFu; };
    // column pair (i, i + 1) as one dword: held in v when i is even, else
    // assembled from two neighbouring dwords (one v_alignbit each)
    auto col_pair = [&](int i) -> uint32_t {
      return (i & 1) ? __builtin_amdgcn_alignbit(v[(i >> 1) + 1], v[i >> 1], 16) : v[i >> 1];
    };
    uint32_t lo = 0, hi = 0;
    if constexpr (sep_horizontal_u16<MODE, K>()) {
      // two outputs per 16-bit-lane multiply-add: outputs (b, b + 1) take
      // column-sum pairs (i, i + 1), an aligned dword of v for even i
      // (every tap when C is even), else one v_alignbit
      auto mac = [](sep_us2 acc, sep_us2 x, sep_us2 wt) {
        if (MODE == kSepSaturate) {
          sep_us2 r;
          asm("v_pk_mad_u16 %0, %1, %2, %3 clamp" : "=v"(r) : "v"(x), "s"(wt), "v"(acc));
          return r;
        }
        return acc + x * wt;
      };
      const uint16_t rnd = MODE == kSepBinomial ? static_cast<uint16_t>(1u << (a.shift - 1)) : 0;
      const uint16_t sh = static_cast<uint16_t>(a.shift);
      uint32_t word[2];
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        sep_us2 acc[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int b = 4 * half + 2 * q;
          sep_us2 s = sep_us2{0, 0};
#pragma unroll
          for (int j = 0; j < K; ++j)
            s = mac(s, __builtin_bit_cast(sep_us2, col_pair(8 * D + b + (j - R) * C)),
                    sep_us2{static_cast<uint16_t>(w[j]), static_cast<uint16_t>(w[j])});
          if (MODE == kSepBinomial)
            s = (s + sep_us2{rnd, rnd}) >> sep_us2{sh, sh};
          else
            s = __builtin_elementwise_min(s, sep_us2{255, 255});
          acc[q] = s;
        }
        word[half] = __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, acc[1]),
                                           __builtin_bit_cast(uint32_t, acc[0]), 0x06040200u);
      }
      lo = word[0];
      hi = word[1];
    } else {
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint32_t acc = 0;
      if (C == 1) {
        // adjacent taps two at a time: v_dot2_u32_u16 (exact, no clamp)
#pragma unroll
        for (int j = 0; j + 1 < K; j += 2)
          acc = __builtin_amdgcn_udot2(__builtin_bit_cast(sep_us2, col_pair(8 * D + b + j - R)),
                                       wpair[j / 2], acc, false);
        acc += __umul24(w[K - 1], col(8 * D + b + K - 1 - R));
      } else {
        // taps C columns apart, two at a time: one v_perm_b32 packs the two
        // column sums (low half first), then v_dot2_u32_u16 as above
#pragma unroll
        for (int j = 0; j + 1 < K; j += 2) {
          const int i0 = 8 * D + b + (j - R) * C, i1 = i0 + C;
          const uint32_t sel = ((i0 & 1) ? 0x0302u : 0x0100u) | (((i1 & 1) ? 0x0706u : 0x0504u) << 16);
          const uint32_t pair = __builtin_amdgcn_perm(v[i1 >> 1], v[i0 >> 1], sel);
          acc = __builtin_amdgcn_udot2(__builtin_bit_cast(sep_us2, pair), wpair[j / 2], acc, false);
        }
        acc += __umul24(w[K - 1], col(8 * D + b + (K - 1 - R) * C));
      }
      uint32_t out;
      if (MODE == kSepBinomial)
        out = (acc + (1u << (a.shift - 1))) >> a.shift;
      else if (MODE == kSepSigma)
        out = (acc + 128u) >> 8;
      else
        out = min(acc, 255u);
      if (b < 4) lo |= (out & 0xFFu) << (8 * b);
      else hi |= (out & 0xFFu) << (8 * (b - 4));
    }
    }
    uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride;
    const int e = e0 + 8 * tx;
    typedef uint2 __attribute__((aligned(1))) u2u;
    if (e + 7 < row_bytes && (reinterpret_cast<uintptr_t>(out + e) & 7) == 0) {
      store_nt(reinterpret_cast<uint2 *>(out + e), make_uint2(lo, hi));
    } else if (e + 7 < row_bytes) {
      *reinterpret_cast<u2u *>(out + e) = make_uint2(lo, hi);  // unaligned rows
    } else {
      store_quad(out, e, row_bytes, lo);
      store_quad(out, e + 4, row_bytes, hi);
    }
  }
}

// One 256-byte x TH-row output tile at tile coordinates (bx, by).  `smem`
// holds separable_smem_bytes<K, TH>() bytes, 16-byte aligned.
template <int MODE, int K, int C, int TH>
__device__ __forceinline__ void separable_tile(const SepArgs &a, int bx, int by,
                                               uint8_t *smem) {
  constexpr int R = K / 2;
  using G = SepWindow<K, TH>;
  uint8_t *win = smem;
  uint16_t *vert = reinterpret_cast<uint16_t *>(smem + separable_win_bytes<K, TH>());

  const int e0 = bx * kTileW;
  const int y0 = by * TH;
  __syncthreads();  // LDS may still hold the previous tile of a batched loop
  stage_window<R, C, TH>(win, a.src, a.src_stride, a.width, a.height, a.border, e0, y0);
  __syncthreads();

  // Vertical pass: each item is one window dword column over two output
  // rows, so its K + 1 window dwords are read and split into 16-bit lanes
  // once for both rows; v_pk_mad_u16 does two lanes per op (with clamp in
  // the saturating mode).
  constexpr int kQuads = G::kCols / 4;
  constexpr int kPairs = TH / 2;
  const uint32_t *win32 = reinterpret_cast<const uint32_t *>(win);
  sep_us2 wy[K];
#pragma unroll
  for (int k = 0; k < K; ++k)
    wy[k] = sep_us2{static_cast<uint16_t>(a.wy[k]), static_cast<uint16_t>(a.wy[k])};
  auto mac = [](sep_us2 acc, sep_us2 x, sep_us2 w) {
    if (MODE == kSepSaturate) {
      sep_us2 r;
      // the taps are uniform: an SGPR operand saves a VGPR copy per use
      asm("v_pk_mad_u16 %0, %1, %2, %3 clamp" : "=v"(r) : "v"(x), "s"(w), "v"(acc));
      return r;
    }
    return acc + x * w;
  };
  for (int i = threadIdx.x; i < kPairs * kQuads; i += kThreads) {
    const int p = i / kQuads, q = i - p * kQuads;  // rows 2p and 2p + 1
    // bytes 0,2 (even) and 1,3 (odd) of each row
    sep_us2 e0 = sep_us2{0, 0}, o0 = sep_us2{0, 0}, e1 = sep_us2{0, 0}, o1 = sep_us2{0, 0};
#pragma unroll
    for (int k = 0; k <= K; ++k) {
      const uint32_t v = win32[(2 * p + k) * kQuads + q];
      const sep_us2 e = __builtin_bit_cast(sep_us2, v & 0x00FF00FFu);
      // bytes 1 and 3 into the low bytes of the halves: one v_perm_b32
      const sep_us2 o = __builtin_bit_cast(sep_us2, __builtin_amdgcn_perm(0u, v, 0x0C030C01u));
      if (k < K) {
        e0 = mac(e0, e, wy[k]);
        o0 = mac(o0, o, wy[k]);
      }
      if (k > 0) {
        e1 = mac(e1, e, wy[k - 1]);
        o1 = mac(o1, o, wy[k - 1]);
      }
    }
    if (MODE == kSepSigma) {
      e0 = (e0 + sep_us2{128, 128}) >> 8;
      o0 = (o0 + sep_us2{128, 128}) >> 8;
      e1 = (e1 + sep_us2{128, 128}) >> 8;
      o1 = (o1 + sep_us2{128, 128}) >> 8;
    }
    // column order 0, 1, 2, 3 = even.x, odd.x, even.y, odd.y
    auto put = [&](int r, sep_us2 even, sep_us2 odd) {
      const uint32_t ev = __builtin_bit_cast(uint32_t, even);
      const uint32_t od = __builtin_bit_cast(uint32_t, odd);
      *reinterpret_cast<uint2 *>(vert + r * G::kCols + 4 * q) =
          make_uint2((ev & 0xFFFFu) | (od << 16), (ev >> 16) | (od & 0xFFFF0000u));
    };
    put(2 * p, e0, o0);
    put(2 * p + 1, e1, o1);
  }
  __syncthreads();

  separable_horizontal<MODE, K, C, TH>(a, vert, e0, y0);
}

// ---------------------------------------------------------------------------
// Blur and downsample (pyrDown): the binomial 5x5 Gaussian at even source
// rows and columns, output ((w+1)/2) x ((h+1)/2).
//
// An output tile is kDownRows rows x 32 * OB bytes, OB = 4 output bytes per
// thread (3 when C == 3), so OB * thread and the tile start are multiples of
// C and every output byte's channel is a compile-time constant.  The source
// window is the radius-2 window of the separable filter over source rows
// [2 y0' - 2, 2 y0' + 18) and source bytes from twice the tile's first output
// byte, staged once; a vertical pass writes [1,4,6,4,1] column sums (u16,
// at most 4080) for the even source rows only, and each thread then reads
// the column sums of its outputs with aligned LDS loads.  One rounding:
// (sum + 128) >> 8, as src/filters/blur_and_downsample_neon.cpp:193-225.
constexpr int kDownRows = 8;

template <int C>
struct Down {
  static constexpr int kOB = C == 3 ? 3 : 4;
  static constexpr int kOutW = 32 * kOB;  // output bytes per tile row
};

// The staged window: 20 source rows (radius 2 around 16).
constexpr int blur_down_win_bytes() {
  return (Window<2>::kRows * Window<2>::kCols + 15) / 16 * 16;
}

constexpr int blur_down_smem_bytes() {
  return blur_down_win_bytes() + 2 * kDownRows * Window<2>::kCols;
}

template <int C>
__device__ __forceinline__ void blur_down_tile(const SepArgs &a, int bx, int by,
                                               uint8_t *smem) {
  using G = Window<2>;
  constexpr int OB = Down<C>::kOB;
  uint8_t *win = smem;
  uint16_t *vert = reinterpret_cast<uint16_t *>(smem + blur_down_win_bytes());
  const int o0 = bx * Down<C>::kOutW;  // first output byte of the tile
  const int y0 = 2 * kDownRows * by;   // first source row of the tile
  __syncthreads();  // LDS may still hold the previous tile of a batched loop
  stage_window<2, C>(win, a.src, a.src_stride, a.width, a.height, a.border, 2 * o0, y0);
  __syncthreads();

  // vertical pass: output row r uses window rows 2r .. 2r + 4
  constexpr int kQuads = G::kCols / 4;
  const uint32_t *win32 = reinterpret_cast<const uint32_t *>(win);
  for (int i = threadIdx.x; i < kDownRows * kQuads; i += kThreads) {
    const int r = i / kQuads, q = i - r * kQuads;
    uint32_t v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = win32[(2 * r + k) * kQuads + q];
    auto lanes = [&](int shift) {
      const uint32_t m = 0x00FF00FFu;
      const sep_us2 x0 = __builtin_bit_cast(sep_us2, (v[0] >> shift) & m);
      const sep_us2 x1 = __builtin_bit_cast(sep_us2, (v[1] >> shift) & m);
      const sep_us2 x2 = __builtin_bit_cast(sep_us2, (v[2] >> shift) & m);
      const sep_us2 x3 = __builtin_bit_cast(sep_us2, (v[3] >> shift) & m);
      const sep_us2 x4 = __builtin_bit_cast(sep_us2, (v[4] >> shift) & m);
      return (x0 + x4) + ((x1 + x3) << 2) + x2 * sep_us2{6, 6};
    };
    const uint32_t ev = __builtin_bit_cast(uint32_t, lanes(0));
    const uint32_t od = __builtin_bit_cast(uint32_t, lanes(8));
    *reinterpret_cast<uint2 *>(vert + r * G::kCols + 4 * q) =
        make_uint2((ev & 0xFFFFu) | (od << 16), (ev >> 16) | (od & 0xFFFF0000u));
  }
  __syncthreads();

  // horizontal pass: thread tx of row r makes output bytes o0 + OB tx + b.
  // Output byte b with channel c = b % C takes the column sums at window
  // u16 index kPad + 2 (OB tx + b) - c + (j - 2) C, j = 0..4; relative to
  // the thread's first loaded index (8 tx for OB = 4, 6 tx + 8 for OB = 3)
  // that is a compile-time offset.
  const int tx = threadIdx.x & 31, r = threadIdx.x >> 5;
  const int yo = kDownRows * by + r;
  const int out_rows = (a.height + 1) / 2;
  const int out_bytes = ((a.width + 1) / 2) * C;
  const int eo = o0 + OB * tx;
  if (yo >= out_rows || eo >= out_bytes) return;
  const uint16_t *row = vert + r * G::kCols;
  constexpr int kWords = OB == 4 ? 16 : 9;
  constexpr int kBase = OB == 4 ? 0 : 8;  // first loaded u16 index - 6 tx
  uint32_t w[kWords];
  if constexpr (OB == 4) {
    const uint4 *p = reinterpret_cast<const uint4 *>(row + 8 * tx);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const uint4 q = p[m];
      w[4 * m] = q.x;
      w[4 * m + 1] = q.y;
      w[4 * m + 2] = q.z;
      w[4 * m + 3] = q.w;
    }
  } else {
    const uint32_t *p = reinterpret_cast<const uint32_t *>(row + 6 * tx + kBase);
#pragma unroll
    for (int m = 0; m < kWords; ++m) w[m] = p[m];
  }
  auto col = [&](int li) -> uint32_t { return (w[li >> 1] >> (16 * (li & 1))) & 0xFFFFu; };
  uint32_t packed = 0;
#pragma unroll
  for (int b = 0; b < OB; ++b) {
    const int c = b % C;
    const int li = G::kPad + 2 * b - c - 2 * C - kBase;
    const uint32_t acc = col(li) + col(li + 4 * C) + ((col(li + C) + col(li + 3 * C)) << 2) +
                         6 * col(li + 2 * C);
    packed |= ((acc + 128u) >> 8) << (8 * b);
  }
  uint8_t *out = a.dst + static_cast<size_t>(yo) * a.dst_stride + eo;
  if (OB == 4 && eo + 4 <= out_bytes && (reinterpret_cast<uintptr_t>(out) & 3) == 0) {
    store_nt(reinterpret_cast<uint32_t *>(out), packed);
  } else {
#pragma unroll
    for (int b = 0; b < OB; ++b)
      if (eo + b < out_bytes) out[b] = static_cast<uint8_t>(packed >> (8 * b));
  }
}

// ---------------------------------------------------------------------------
// Median: exact selection on pairs of horizontally adjacent outputs packed
// in 16-bit halves (v_pk_min_u16 / v_pk_max_u16).  Any exact selection
// returns the reference sorting network's value.
typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void cmpex(us2 &lo, us2 &hi) {
  const us2 a = lo;
  lo = __builtin_elementwise_min(a, hi);
  hi = __builtin_elementwise_max(a, hi);
}

// Batcher's odd-even merge sort on N elements, generated at compile time
// (comparators that would touch index >= N compare against +infinity and
// are left out).  Applied fully unrolled; comparators whose results are
// never used are removed by the compiler, which turns it into a selection
// network for the ranks that are read.
struct CmpPair {
  int a, b;
};

template <int N>
struct BatcherNet {
  template <class F>
  static constexpr void visit(F f) {
    for (int p = 1; p < N; p <<= 1)
      for (int k = p; k >= 1; k >>= 1)
        for (int j = k % p; j + k < N; j += 2 * k)
          for (int i = 0; i < k; ++i)
            if (i + j + k < N && (i + j) / (2 * p) == (i + j + k) / (2 * p)) f(i + j, i + j + k);
  }
  static constexpr int count() {
    int c = 0;
    visit([&](int, int) { ++c; });
    return c;
  }
  static constexpr int kCount = count();
  struct Table {
    CmpPair p[kCount > 0 ? kCount : 1];
  };
  static constexpr Table make() {
    Table t{};
    int c = 0;
    visit([&](int a, int b) { t.p[c++] = CmpPair{a, b}; });
    return t;
  }
  static constexpr Table kTable = make();
};

template <int N>
__device__ __forceinline__ void batcher_sort(us2 *a) {
#pragma unroll
  for (int c = 0; c < BatcherNet<N>::kCount; ++c)
    cmpex(a[BatcherNet<N>::kTable.p[c].a], a[BatcherNet<N>::kTable.p[c].b]);
}

// Batcher's odd-even merge of two sorted runs of N and M values, as
// compare-exchanges on N + M registers (run A in registers 0..N-1, run B in
// N..N+M-1).  Both runs are padded with +infinity to a power of two P and
// the network is traced at compile time: a comparator that meets a padding
// slot only moves a register (or nothing), so it is not emitted, and
// rank[k] names the register holding the k-th smallest value at the end.
template <int N, int M>
struct MergeNet {
  static constexpr int pow2_at_least(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
  }
  static constexpr int P = pow2_at_least(N > M ? N : M);
  struct Table {
    CmpPair c[16 * P];  // a merge of 2P values has at most P * (log2(P) + 1) + 1
    int count;
    int rank[N + M];
  };
  struct Trace {
    int pos[2 * P];  // register at each position, -1 for padding
    Table t;
    constexpr void cmp(int i, int j) {
      const int a = pos[i], b = pos[j];
      if (a >= 0 && b >= 0) {
        t.c[t.count++] = CmpPair{a, b};
      } else if (a < 0 && b >= 0) {
        pos[i] = b;
        pos[j] = -1;
      }
    }
    constexpr void merge(int lo, int n, int r) {
      const int m = 2 * r;
      if (m < n) {
        merge(lo, n, m);
        merge(lo + r, n, m);
        for (int i = lo + r; i + r < lo + n; i += m) cmp(i, i + r);
      } else {
        cmp(lo, lo + r);
      }
    }
  };
  static constexpr Table make() {
    Trace tr{};
    for (int i = 0; i < P; ++i) {
      tr.pos[i] = i < N ? i : -1;
      tr.pos[P + i] = i < M ? N + i : -1;
    }
    tr.t.count = 0;
    tr.merge(0, 2 * P, 1);
    for (int k = 0; k < N + M; ++k) tr.t.rank[k] = tr.pos[k];
    return tr.t;
  }
  static constexpr Table kTable = make();
};

// out[0..N+M) = merge of sorted a[0..N) and sorted b[0..M).
template <int N, int M>
__device__ __forceinline__ void merge_runs(const us2 *a, const us2 *b, us2 *out) {
  using Net = MergeNet<N, M>;
  us2 v[N + M];
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = a[i];
#pragma unroll
  for (int i = 0; i < M; ++i) v[N + i] = b[i];
#pragma unroll
  for (int c = 0; c < Net::kTable.count; ++c) cmpex(v[Net::kTable.c[c].a], v[Net::kTable.c[c].b]);
#pragma unroll
  for (int k = 0; k < N + M; ++k) out[k] = v[Net::kTable.rank[k]];
}

// Median of each K x K window, computed for four vertically adjacent
// outputs at once.  Windows 0 and 1 share window rows 1..K-1, windows 2 and
// 3 share rows 3..K+1, and all four share rows 3..K-1.  A shared value of
// rank r among the K(K-1) values two adjacent windows share has rank
// r..r+K in either window, so only shared ranks m-K-1..m-1 (m = (K*K+1)/2,
// 0-based) can be a window's median; lower ones are always below it, higher
// ones above.  Each window's median is therefore the (K+1)-th smallest of
// those K+1 shared values and its own sorted row:
//   min over i = 0..K of max(A[K-i], B[i-1])   (B[-1] = -infinity).
// Every window row is sorted once (K-input network); the rows common to a
// pair of windows are combined with merge networks, the rows common to all
// four windows only once.  For 5x5 this is 167 packed min/max operations
// per two outputs instead of 224 for pairs alone.
template <int K>
constexpr int median_smem_bytes() {
  return Window<K / 2>::kRows * Window<K / 2>::kCols;
}

// floor(v / 4), also for negative v.
constexpr int floor_div4(int v) { return v >= 0 ? v / 4 : -((3 - v) / 4); }

// Bytes b and b + 1 of the little-endian dword run d, zero-extended into the
// two 16-bit halves (one v_perm_b32; b is a compile-time constant).
template <int ND>
__device__ __forceinline__ us2 byte_pair(const uint32_t (&d)[ND], int b) {
  const int q = b >> 2, s = b & 3;
  const uint32_t sel = s | (0x0cu << 8) | ((s + 1u) << 16) | (0x0cu << 24);
  const uint32_t hi = s == 3 ? d[q + 1] : d[q];
  return __builtin_bit_cast(us2, __builtin_amdgcn_perm(hi, d[q], sel));
}

// (K+1)-th smallest (0-based K) of the K + 1 shared candidates c and the
// window's own sorted row r.
template <int K>
__device__ __forceinline__ us2 median_select(const us2 *c, const us2 *r) {
  us2 m = c[K];
#pragma unroll
  for (int i = 1; i <= K; ++i)
    m = __builtin_elementwise_min(m, __builtin_elementwise_max(c[K - i], r[i - 1]));
  return m;
}

// Sorted union of rows [first, first + n) of the sorted rows s (n >= 1),
// merged pairwise.
template <int K, int N>
struct RowMerge {
  static __device__ __forceinline__ void run(const us2 (*s)[K], us2 *out) {
    constexpr int H = N / 2;
    us2 lo[H * K], hi[(N - H) * K];
    RowMerge<K, H>::run(s, lo);
    RowMerge<K, N - H>::run(s + H, hi);
    merge_runs<H * K, (N - H) * K>(lo, hi, out);
  }
};
template <int K>
struct RowMerge<K, 1> {
  static __device__ __forceinline__ void run(const us2 (*s)[K], us2 *out) {
#pragma unroll
    for (int j = 0; j < K; ++j) out[j] = s[0][j];
  }
};

}  // namespace filt
}  // namespace kcv
#include "median5_net.h"
namespace kcv {
namespace filt {

// 5x5 median tile for one and two channels: each 16-bit half of a thread's
// us2 lanes takes a 2-column x 4-row block of same-channel outputs (one
// channel: bytes 0-1 and 2-3 of the thread's dword; two channels: channel
// 0 in bytes 0 and 2, channel 1 in bytes 1 and 3), so the eight windows of
// a block share one 6 x 8 footprint and median5_net_2x4 (558 min/max for
// 16 outputs, against 668 for four columns of 4-row windows) finds all of
// them.  Footprint values are paired from aligned LDS dwords by v_perm_b32.
template <int C>
__device__ __forceinline__ void median5_tile_2x4(const MedianArgs &a, int bx, int by,
                                                 uint8_t *smem) {
  static_assert(C == 1 || C == 2, "blocks of same-channel pairs in one dword");
  constexpr int R = 2;
  using G = Window<R>;
  uint8_t *win = smem;
  const int e0 = bx * kTileW;
  const int y0 = by * kTileH;
  __syncthreads();  // LDS may still hold the previous tile of a batched loop
  stage_window<R, C>(win, a.src, a.src_stride, a.width, a.height, a.border, e0, y0);
  __syncthreads();

  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int r = 4 * ty;  // tile rows r .. r + 3; footprint rows r .. r + 7
  if (y0 + r >= a.height) return;
  // dwords holding bytes -4 .. 7 around the thread's first output byte
  const uint32_t *w0 =
      reinterpret_cast<const uint32_t *>(win + r * G::kCols + G::kPad) + tx - 1;
  us2 f[8][6];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    uint32_t d[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) d[q] = w0[i * (G::kCols / 4) + q];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      // byte positions in d (0 .. 11) of footprint column c in both halves
      const int px = C == 1 ? 2 + c : 2 * c;
      const int py = C == 1 ? 4 + c : 2 * c + 1;
      const int q = px >> 2;
      const uint32_t sel = static_cast<uint32_t>(px & 3) | (0x0cu << 8) |
                           (static_cast<uint32_t>(py - 4 * q) << 16) | (0x0cu << 24);
      f[i][c] = __builtin_bit_cast(us2, __builtin_amdgcn_perm(d[q < 2 ? q + 1 : 2], d[q], sel));
    }
  }
  us2 out[4][2];
  median5_net_2x4(f, out);
  const int row_bytes = a.width * C;
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const int y = y0 + r + h;
    const uint32_t b0 = C == 1 ? out[h][0].x : out[h][0].x;
    const uint32_t b1 = C == 1 ? out[h][1].x : out[h][0].y;
    const uint32_t b2 = C == 1 ? out[h][0].y : out[h][1].x;
    const uint32_t b3 = out[h][1].y;
    if (y < a.height)
      store_quad(a.dst + static_cast<size_t>(y) * a.dst_stride, e0 + 4 * tx, row_bytes,
                 b0 | (b1 << 8) | (b2 << 16) | (b3 << 24));
  }
}

// One 256-byte x 16-row median tile of a C-channel image at tile
// coordinates (bx, by).  Each thread owns 4 output bytes (two us2 lanes) of
// 4 consecutive rows.  The taps of a window row are read from LDS as
// aligned dwords (unaligned 8/16-bit LDS reads are several times slower)
// and paired with v_perm_b32.
template <int K, int C>
__device__ __forceinline__ void median_tile_c(const MedianArgs &a, int bx, int by,
                                              uint8_t *smem) {
  if constexpr (K == 5 && C <= 2) {
    median5_tile_2x4<C>(a, bx, by, smem);
    return;
  }
  constexpr int R = K / 2;
  constexpr int M = (K * K + 1) / 2;     // median rank (1-based)
  constexpr int LO = M - K - 1;          // first kept shared rank (0-based)
  constexpr int NR = K + 3;              // window rows of four outputs
  // dwords covering bytes [-R*C, 3 + R*C] around the thread's first byte
  constexpr int DLO = floor_div4(-R * C), ND = floor_div4(3 + R * C) - DLO + 1;
  using G = Window<R>;
  static_assert(-4 * DLO <= G::kPad && 4 * (DLO + ND) <= 4 + G::kPad,
                "taps stay inside the staged window");
  uint8_t *win = smem;

  const int e0 = bx * kTileW;
  const int y0 = by * kTileH;
  __syncthreads();  // LDS may still hold the previous tile of a batched loop
  stage_window<R, C>(win, a.src, a.src_stride, a.width, a.height, a.border, e0, y0);
  __syncthreads();

  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int r = 4 * ty;  // tile rows r .. r + 3
  if (y0 + r >= a.height) return;
  const uint32_t *w0 =
      reinterpret_cast<const uint32_t *>(win + r * G::kCols + G::kPad) + tx + DLO;
  uint32_t packed[4] = {0, 0, 0, 0};
#pragma unroll
  for (int pair = 0; pair < 2; ++pair) {
    // sorted window rows 0 .. K+2 (row i of output 0's window is tile row r + i - R)
    us2 s[NR][K];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      uint32_t d[ND];
#pragma unroll
      for (int q = 0; q < ND; ++q) d[q] = w0[i * (G::kCols / 4) + q];
#pragma unroll
      for (int j = 0; j < K; ++j) s[i][j] = byte_pair(d, 2 * pair + (j - R) * C - 4 * DLO);
      batcher_sort<K>(s[i]);
    }
    // rows 1..K-1 (outputs 0, 1) and 3..K+1 (outputs 2, 3)
    us2 ca[K * (K - 1)], cb[K * (K - 1)];
    if constexpr (K == 3) {
      merge_runs<K, K>(s[1], s[2], ca);
      merge_runs<K, K>(s[3], s[4], cb);
    } else {
      us2 mid[K * (K - 3)], e1[2 * K];
      RowMerge<K, K - 3>::run(s + 3, mid);
      merge_runs<K, K>(s[1], s[2], e1);
      merge_runs<2 * K, K * (K - 3)>(e1, mid, ca);
      merge_runs<K, K>(s[K], s[K + 1], e1);
      merge_runs<K * (K - 3), 2 * K>(mid, e1, cb);
    }
    const us2 out[4] = {median_select<K>(ca + LO, s[0]), median_select<K>(ca + LO, s[K]),
                        median_select<K>(cb + LO, s[2]), median_select<K>(cb + LO, s[K + 2])};
#pragma unroll
    for (int h = 0; h < 4; ++h)
      packed[h] |= (static_cast<uint32_t>(out[h].x) | (static_cast<uint32_t>(out[h].y) << 8))
                   << (16 * pair);
  }
  const int row_bytes = a.width * C;
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const int y = y0 + r + h;
    if (y < a.height)
      store_quad(a.dst + static_cast<size_t>(y) * a.dst_stride, e0 + 4 * tx, row_bytes,
                 packed[h]);
  }
}

}  // namespace filt
}  // namespace kcv

#endif  // KCV_FILTERS_BODY_H
