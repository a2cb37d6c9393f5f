// Per-image launches of the tiled 2-D filters (bodies in filters_body.h).

#include <type_traits>

#include "filters_body.h"
#include "prepare.h"

namespace kcv {
namespace {

using namespace filt;

// Tile (bx, by) of this workgroup: XCD-ordered over a 1-D grid of
// tiles_x * tiles_y workgroups, or the 2-D grid position when tiles_x == 0.
__device__ __forceinline__ void tile_of(int tiles_x, int &bx, int &by) {
  if (tiles_x == 0) {
    bx = blockIdx.x;
    by = blockIdx.y;
    return;
  }
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  by = t / tiles_x;
  bx = t - by * tiles_x;
}

template <int MODE, int K, int C>
__global__ __launch_bounds__(kThreads) void separable_kernel(SepArgs a, int tiles_x) {
  __shared__ __align__(16) uint8_t smem[separable_smem_bytes<K, sep_tile_rows(K)>()];
  int bx, by;
  tile_of(tiles_x, bx, by);
  separable_tile<MODE, K, C, sep_tile_rows(K)>(a, bx, by, smem);
}

// One kernel per channel count: the tile body is specialised on it.
template <int MODE, int K>
void launch_channels(const SepArgs &a, dim3 grid, int tx, hipStream_t s) {
  switch (a.channels) {
    case 1: separable_kernel<MODE, K, 1><<<grid, kThreads, 0, s>>>(a, tx); break;
    case 2: separable_kernel<MODE, K, 2><<<grid, kThreads, 0, s>>>(a, tx); break;
    case 3: separable_kernel<MODE, K, 3><<<grid, kThreads, 0, s>>>(a, tx); break;
    default: separable_kernel<MODE, K, 4><<<grid, kThreads, 0, s>>>(a, tx); break;
  }
}

template <int K, int C>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(K == 5 && C <= 2 ? 8 : 1))) void median_kernel(MedianArgs a, int tiles_x) {
  __shared__ __align__(16) uint8_t smem[median_smem_bytes<K>()];
  int bx, by;
  tile_of(tiles_x, bx, by);
  median_tile_c<K, C>(a, bx, by, smem);
}

template <int C>
__global__ __launch_bounds__(kThreads) void blur_down_kernel(SepArgs a, int tiles_x) {
  __shared__ __align__(16) uint8_t smem[blur_down_smem_bytes()];
  int bx, by;
  tile_of(tiles_x, bx, by);
  blur_down_tile<C>(a, bx, by, smem);
}

// ---------------------------------------------------------------------------
// Gaussian blur of any odd size (src/filters/gaussian_blur_arbitrary_neon.cpp):
// REPLICATE border, symmetric u8 taps summing to 256, vertical pass rounded
// to u8, then the horizontal pass rounded again.
constexpr int kArbRows = 16;        // output rows per thread in the vertical pass
constexpr int kArbRowBytes = 1024;  // output bytes per workgroup in the horizontal pass

// Vertical pass: thread = one dword column (4 bytes) x kArbRows output rows.
// Source rows from y0 - R are taken in groups of four: the four row dwords of
// a group are transposed by v_perm_b32 into one dword per column holding
// that column's four rows, and each output row r adds v_dot4_u32_u8 of it
// with a weight dword.  Relative to y0 - R, group g byte b is row offset
// 4g + b, whose tap for output row r = 4q + s is t[4(g - q) + b - s]: the
// weight is phase[s][g - q], built on the host (zero bytes outside the
// kernel), and uniform, so it is read with scalar loads.  The sums are exact
// in 32 bits and round to u8 as the reference's vertical pass does.
constexpr int kArbVDwords = (3 + 254) / 4 + 1;  // row groups per output row at K = 255

struct GaussArbVertical {
  GaussArbArgs a;
  uint32_t phase[4 * kArbVDwords];
};

__global__ __launch_bounds__(kThreads) void gauss_arb_vertical_kernel(GaussArbVertical v) {
  const GaussArbArgs &a = v.a;
  const int row_bytes = a.width * a.channels;
  const int e = 4 * (blockIdx.x * kThreads + threadIdx.x);
  if (e >= row_bytes) return;
  const int K = a.ksize, R = K / 2;
  const int nd = (K + 2) / 4 + 1;              // weight dwords per phase
  const int ng = (kArbRows + K + 2) / 4;       // groups over rows y0 - R .. y0 + kArbRows - 1 + R
  const bool vec = e + 4 <= row_bytes &&
                   ((reinterpret_cast<uintptr_t>(a.src) | a.src_stride) & 3) == 0;
  for (int y0 = blockIdx.y * kArbRows; y0 < a.height; y0 += gridDim.y * kArbRows) {
    uint32_t acc[kArbRows][4];
#pragma unroll
    for (int r = 0; r < kArbRows; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[r][c] = 0;
    for (int g = 0; g < ng; ++g) {
      uint32_t d[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int j = min(max(y0 - R + 4 * g + k, 0), a.height - 1);
        const uint8_t *row = a.src + static_cast<size_t>(j) * a.src_stride;
        if (vec) {
          d[k] = *reinterpret_cast<const uint32_t *>(row + e);
        } else {
          d[k] = 0;
          for (int b = 0; b < 4 && e + b < row_bytes; ++b)
            d[k] |= static_cast<uint32_t>(row[e + b]) << (8 * b);
        }
      }
      // 4x4 byte transpose: col[c] = (d0.c, d1.c, d2.c, d3.c)
      const uint32_t lo01 = __builtin_amdgcn_perm(d[1], d[0], 0x05010400u);
      const uint32_t hi01 = __builtin_amdgcn_perm(d[1], d[0], 0x07030602u);
      const uint32_t lo23 = __builtin_amdgcn_perm(d[3], d[2], 0x05010400u);
      const uint32_t hi23 = __builtin_amdgcn_perm(d[3], d[2], 0x07030602u);
      const uint32_t col[4] = {__builtin_amdgcn_perm(lo23, lo01, 0x05040100u),
                               __builtin_amdgcn_perm(lo23, lo01, 0x07060302u),
                               __builtin_amdgcn_perm(hi23, hi01, 0x05040100u),
                               __builtin_amdgcn_perm(hi23, hi01, 0x07060302u)};
#pragma unroll
      for (int r = 0; r < kArbRows; ++r) {
        const int m = g - r / 4;
        if (m < 0 || m >= nd) continue;
        const uint32_t w = v.phase[(r & 3) * kArbVDwords + m];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = __builtin_amdgcn_udot4(col[c], w, acc[r][c], false);
      }
    }
#pragma unroll
    for (int r = 0; r < kArbRows; ++r) {
      const int y = y0 + r;
      if (y >= a.height) break;
      const uint32_t out = ((acc[r][0] + 128) >> 8) | (((acc[r][1] + 128) >> 8) << 8) |
                           (((acc[r][2] + 128) >> 8) << 16) | (((acc[r][3] + 128) >> 8) << 24);
      uint8_t *o = a.tmp + static_cast<size_t>(y) * a.dst_stride + e;
      if (vec && ((reinterpret_cast<uintptr_t>(o) & 3) == 0)) {
        *reinterpret_cast<uint32_t *>(o) = out;
      } else {
        for (int b = 0; b < 4 && e + b < row_bytes; ++b) o[b] = static_cast<uint8_t>(out >> (8 * b));
      }
    }
  }
}

// Horizontal pass: a workgroup takes kArbRowBytes output bytes of one row
// and stages the intermediate span they read (border pixels replicated) in
// LDS; each thread makes 4 consecutive output bytes o0 + p from the aligned
// span dwords at o0 with v_dot4_u32_u8.  Output o0 + p sums taps[i] *
// span[o0 + p + i * C]: against span dword m, byte b, that is the tap
// i = (4m + b - p) / C where 4m + b - p is a non-negative multiple of C below
// K * C, so each phase p has its own weight dwords (zero bytes elsewhere),
// built once per workgroup.  The sums are exact in 32 bits.
constexpr int kArbPhaseDwords = (3 + 254 * 4) / 4 + 1;  // K = 255, C = 4

__global__ __launch_bounds__(kThreads) void gauss_arb_horizontal_kernel(GaussArbArgs a) {
  __shared__ uint8_t taps[256];
  __shared__ uint32_t wq[4 * kArbPhaseDwords];
  // span dwords read past n (up to 4 * kArbPhaseDwords bytes after a thread's
  // first) meet zero weights only
  __shared__ __align__(16) uint8_t span[kArbRowBytes + 4 * kArbPhaseDwords];
  taps[threadIdx.x] = a.taps[threadIdx.x];
  const int C = a.channels, K = a.ksize, R = K / 2;
  const int nd = (3 + (K - 1) * C) / 4 + 1;  // span dwords per output
  __syncthreads();
  for (int i = threadIdx.x; i < 4 * nd; i += kThreads) {
    const int p = i / nd, m = i - p * nd;
    uint32_t w = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int d = 4 * m + b - p;
      if (d >= 0 && d % C == 0 && d / C < K) w |= static_cast<uint32_t>(taps[d / C]) << (8 * b);
    }
    wq[i] = w;
  }
  const int row_bytes = a.width * C;
  const int e0 = blockIdx.x * kArbRowBytes;
  const int n = min(kArbRowBytes, row_bytes - e0) + 2 * R * C;  // span bytes
  const int o0 = 4 * threadIdx.x;
  for (int y = blockIdx.y; y < a.height; y += gridDim.y) {
    const uint8_t *row = a.tmp + static_cast<size_t>(y) * a.dst_stride;
    __syncthreads();  // the previous row's span is no longer read (and wq is built)
    for (int b = threadIdx.x; b < n; b += kThreads) {
      const int eb = e0 - R * C + b;  // byte position, may lie outside the row
      const int x = eb >= 0 ? eb / C : -((C - 1 - eb) / C);
      const int c = eb - x * C;
      span[b] = row[min(max(x, 0), a.width - 1) * C + c];
    }
    __syncthreads();
    if (e0 + o0 >= row_bytes) continue;
    const uint32_t *s32 = reinterpret_cast<const uint32_t *>(span) + threadIdx.x;
    uint32_t acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
    for (int m = 0; m < nd; ++m) {
      const uint32_t d = s32[m];
      acc0 = __builtin_amdgcn_udot4(d, wq[m], acc0, false);
      acc1 = __builtin_amdgcn_udot4(d, wq[nd + m], acc1, false);
      acc2 = __builtin_amdgcn_udot4(d, wq[2 * nd + m], acc2, false);
      acc3 = __builtin_amdgcn_udot4(d, wq[3 * nd + m], acc3, false);
    }
    const uint32_t out = ((acc0 + 128) >> 8) | (((acc1 + 128) >> 8) << 8) |
                         (((acc2 + 128) >> 8) << 16) | (((acc3 + 128) >> 8) << 24);
    store_quad(a.dst + static_cast<size_t>(y) * a.dst_stride, e0 + o0, row_bytes, out);
  }
}

// ---------------------------------------------------------------------------
// Median of larger windows (the sizes the reference serves with histograms,
// src/filters/median_blur_{small,large}_hist_neon.cpp).  Any exact selection
// gives the same value; here each lane keeps the histogram of its own
// window and slides it down a strip of output rows (Huang's update): per
// output row it adds the window's new bottom row and removes its old top
// row (2K taps) and moves the median to the new rank.
//
// One wave per workgroup, one output byte per lane.  A lane's 256 bins are
// 16-bit counts packed two to a dword (bins 2i and 2i + 1 in dword i, low
// half first; counts reach K*K <= 65025) and interleaved by lane (dword i of
// lane l at i * 64 + l), so every LDS access of a wave hits 64 different
// banks whatever the pixel values: 32 KiB per wave.  Adds and removes are
// fire-and-forget LDS atomics (ds_add_u32).  `lt` counts the window's values
// below the current median m; after each update m moves down while more
// than (K*K-1)/2 values lie below it and up while no more than that lie at
// or below it, reading the eight bins m-4 .. m+3 in one round (five dwords)
// and stepping through them in registers.  A strip starts from a full K x K
// histogram, so strips are at least K rows tall (the start costs K*K taps,
// the walk 2K per row).
constexpr int kHistLanes = 64;
constexpr int kHistPipeTaps = 32;  // widest window row loaded one output row ahead

// Largest window served by the merge-network tile; larger ones slide
// histograms.
#ifndef KCV_MEDIAN_NETWORK_MAX
#define KCV_MEDIAN_NETWORK_MAX 11
#endif
constexpr int kMedianNetworkMax = KCV_MEDIAN_NETWORK_MAX;

// Rows per strip of the histogram median: enough workgroups for about four
// rounds over the GPU (five 32 KiB histograms fit one CU's LDS), but at
// least K rows (a strip's start costs K*K taps, each further row 2K) and
// few enough strips for a grid dimension.
int median_hist_strip(int blocks_x, int height, int k, int cus) {
  const int64_t want = int64_t{4} * 5 * std::max(cus, 1);
  int64_t strip = (int64_t{blocks_x} * height + want - 1) / want;
  strip = std::max<int64_t>(strip, k);
  strip = std::max<int64_t>(strip, (height + 65534) / 65535);
  return static_cast<int>(std::min<int64_t>(strip, height));
}

template <int C>
__global__ __launch_bounds__(kHistLanes) void median_hist_kernel(MedianArgs a, int strip) {
  __shared__ uint32_t hist[128 * kHistLanes];
  const int lane = threadIdx.x;
  const int row_bytes = a.width * C;
  const int e = blockIdx.x * kHistLanes + lane;
  const bool active = e < row_bytes;
  const int ec = active ? e : row_bytes - 1;  // idle lanes shadow the last byte
  const int x = ec / C, c = ec - x * C;
  const int K = a.ksize, R = K / 2;
  const int half = (K * K - 1) / 2;
  const ColMap cm = col_map(a.border, a.width);
  const bool interior = x - R >= 0 && x + R < a.width;
  uint32_t *h = hist + lane;
  for (int i = 0; i < 128; ++i) h[i * kHistLanes] = 0;
  int m = 0, lt = 0;
  // One tap value v added (SIGN 1) or removed (SIGN -1) unless `used` (the
  // same for every lane) is false.
  auto count = [&](uint32_t v, bool used, auto sign_tag) {
    constexpr int SIGN = decltype(sign_tag)::value;
    lt += used && v < static_cast<uint32_t>(m) ? SIGN : 0;
    if (used) {
      uint32_t *p = h + (v >> 1) * kHistLanes;
      const uint32_t inc = 1u << (16 * (v & 1));
      atomicAdd(p, SIGN > 0 ? inc : 0u - inc);
    }
  };
  // Column offsets of the taps: direct for interior lanes, through the
  // border map (as selects, no branches) near the left and right edges.
  auto tap_offset = [&](int j) -> int {
    const int xx = x + j - R;
    const int l = cm.sl * xx + cm.bl, r = cm.sr * xx + cm.br;
    const int mapped = min(max(xx < 0 ? l : (xx >= a.width ? r : xx), 0), a.width - 1);
    return (interior ? xx : mapped) * C;
  };
  // Taps of window row yy added or removed.  One-channel interior lanes read
  // their K contiguous bytes as dwords (the last dword is the window's last
  // four bytes, its bytes already counted skipped); other lanes read bytes.
  // Loads go out in groups (4 dwords or 16 bytes) before their LDS atomics;
  // a group's unused slots re-read the last tap and are not counted.
  auto update = [&](int yy, auto sign_tag) {
    const uint8_t *row =
        a.src + static_cast<size_t>(map_index(yy, a.height, a.border)) * a.src_stride + c;
    if (C == 1 && interior) {
      typedef uint32_t __attribute__((aligned(1))) u32u;
      const uint8_t *p = row + x - R;
      const int nd = (K + 3) / 4;
      for (int t0 = 0; t0 < nd; t0 += 4) {
        uint32_t d[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
          d[t] = *reinterpret_cast<const u32u *>(p + min(4 * min(t0 + t, nd - 1), K - 4));
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int start = min(4 * (t0 + t), K - 4);
#pragma unroll
          for (int b = 0; b < 4; ++b)
            count((d[t] >> (8 * b)) & 0xFFu, t0 + t < nd && start + b >= 4 * (t0 + t),
                  sign_tag);
        }
      }
    } else {
      for (int j0 = 0; j0 < K; j0 += 16) {
        uint32_t v[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) v[t] = row[tap_offset(min(j0 + t, K - 1))];
#pragma unroll
        for (int t = 0; t < 16; ++t) count(v[t], j0 + t < K, sign_tag);
      }
    }
  };
  using Add = std::integral_constant<int, 1>;
  using Remove = std::integral_constant<int, -1>;
  const int y0 = blockIdx.y * strip;
  const int y1 = min(y0 + strip, a.height);
  for (int i = -R; i <= R; ++i) update(y0 + i, Add{});
  // bins m-4 .. m+3 (those outside 0..255 are read clamped and never used)
  auto bins8 = [&](uint32_t (&b)[8]) {
    const int w0 = (m - 4) >> 1;  // arithmetic shift: floor
    uint32_t w[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) w[i] = h[min(max(w0 + i, 0), 127) * kHistLanes];
    const uint32_t sh = static_cast<uint32_t>(m & 1) * 16;  // (m - 4) odd: start high
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t pair = __builtin_amdgcn_alignbit(w[i + 1], w[i], sh);
      b[2 * i] = pair & 0xFFFFu;
      b[2 * i + 1] = pair >> 16;
    }
  };
  // moves m to the median of the current window
  auto search = [&]() {
    for (bool again = true; again;) {
      uint32_t b[8];
      bins8(b);
      const int m0 = m;
      // down while more than `half` values lie below m (b[3] is bin m0-1)
#pragma unroll
      for (int t = 3; t >= 0; --t) {
        if (lt > half) {
          --m;
          lt -= static_cast<int>(b[t]);
        }
      }
      // up (only when m did not move down) while no more than `half` values
      // lie at or below m (b[4] is bin m0)
      bool up = m == m0;
#pragma unroll
      for (int t = 4; t < 8; ++t) {
        if (up && lt + static_cast<int>(b[t]) <= half) {
          lt += static_cast<int>(b[t]);
          ++m;
        } else {
          up = false;
        }
      }
      again = lt > half || up;  // ran past the eight bins read
    }
  };
  auto store = [&](int y) {
    if (active) a.dst[static_cast<size_t>(y) * a.dst_stride + e] = static_cast<uint8_t>(m);
  };
  search();
  store(y0);
  if (K <= kHistPipeTaps) {
    // Windows of up to kHistPipeTaps columns: the taps of the next output
    // row's new and old window rows are loaded into registers while the
    // current row's atomics and median search run (two register sets used
    // in turn), so the loads' latency is hidden behind the LDS work.
    auto fetch = [&](int yy, uint32_t (&t)[kHistPipeTaps]) {
      const uint8_t *row =
          a.src + static_cast<size_t>(map_index(yy, a.height, a.border)) * a.src_stride + c;
      if (C == 1 && interior) {
        typedef uint32_t __attribute__((aligned(1))) u32u;
        const uint8_t *p = row + x - R;
        const int nd = (K + 3) / 4;
#pragma unroll
        for (int i = 0; i < kHistPipeTaps / 4; ++i)
          t[i] = *reinterpret_cast<const u32u *>(p + min(4 * min(i, nd - 1), K - 4));
      } else {
#pragma unroll
        for (int i = 0; i < kHistPipeTaps; ++i) t[i] = row[tap_offset(min(i, K - 1))];
      }
    };
    auto apply = [&](const uint32_t (&t)[kHistPipeTaps], auto sign_tag) {
      if (C == 1 && interior) {
        const int nd = (K + 3) / 4;
#pragma unroll
        for (int i = 0; i < kHistPipeTaps / 4; ++i) {
          const int start = min(4 * i, K - 4);
#pragma unroll
          for (int b = 0; b < 4; ++b)
            count((t[i] >> (8 * b)) & 0xFFu, i < nd && start + b >= 4 * i, sign_tag);
        }
      } else {
#pragma unroll
        for (int i = 0; i < kHistPipeTaps; ++i) count(t[i], i < K, sign_tag);
      }
    };
    uint32_t add0[kHistPipeTaps], rem0[kHistPipeTaps], add1[kHistPipeTaps],
        rem1[kHistPipeTaps];
    // output row y: apply (add, rem), prefetch row y + 1's into (nadd, nrem)
    auto step = [&](int y, const uint32_t (&add)[kHistPipeTaps],
                    const uint32_t (&rem)[kHistPipeTaps], uint32_t (&nadd)[kHistPipeTaps],
                    uint32_t (&nrem)[kHistPipeTaps]) {
      if (y + 1 < y1) {
        fetch(y + 1 + R, nadd);
        fetch(y - R, nrem);
      }
      apply(add, Add{});
      apply(rem, Remove{});
      search();
      store(y);
    };
    if (y0 + 1 < y1) {
      fetch(y0 + 1 + R, add0);
      fetch(y0 - R, rem0);
    }
    for (int y = y0 + 1; y < y1; y += 2) {
      step(y, add0, rem0, add1, rem1);
      if (y + 1 >= y1) break;
      step(y + 1, add1, rem1, add0, rem0);
    }
  } else {
    for (int y = y0 + 1; y < y1; ++y) {
      update(y + R, Add{});
      update(y - R - 1, Remove{});
      search();
      store(y);
    }
  }
}

// Grid and tiles_x argument for an image: a 1-D XCD-ordered grid unless the
// tile count is very large.
struct TileGrid {
  dim3 grid;
  int tiles_x;
};

TileGrid tile_grid(int width, int height, int channels, int tile_rows) {
  const int tx = (width * channels + kTileW - 1) / kTileW;
  const int ty = (height + tile_rows - 1) / tile_rows;
  if (int64_t{tx} * ty <= kMaxLinearTiles) return {dim3(tx * ty), tx};
  return {dim3(tx, ty), 0};
}

template <int MODE>
hipError_t launch_mode(const SepArgs &a, hipStream_t s) {
  const TileGrid g = tile_grid(a.width, a.height, a.channels, sep_tile_rows(a.ksize));
  const dim3 grid = g.grid;
  const int tx = g.tiles_x;
  switch (a.ksize) {
    case 3: launch_channels<MODE, 3>(a, grid, tx, s); break;
    case 5: launch_channels<MODE, 5>(a, grid, tx, s); break;
    case 7: launch_channels<MODE, 7>(a, grid, tx, s); break;
    case 9:
      if constexpr (MODE != kSepSaturate) {
        launch_channels<MODE, 9>(a, grid, tx, s);
        break;
      }
      return hipErrorInvalidValue;
    case 15:
      if constexpr (MODE == kSepSigma) {
        launch_channels<MODE, 15>(a, grid, tx, s);
        break;
      }
      return hipErrorInvalidValue;
    case 21:
      if constexpr (MODE == kSepSigma) {
        launch_channels<MODE, 21>(a, grid, tx, s);
        break;
      }
      return hipErrorInvalidValue;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_separable(const SepArgs &a, hipStream_t stream) {
  if (a.width <= 0 || a.height <= 0 || a.channels <= 0) return hipSuccess;
  switch (a.mode) {
    case kSepBinomial: return launch_mode<kSepBinomial>(a, stream);
    case kSepSigma: return launch_mode<kSepSigma>(a, stream);
    default: return launch_mode<kSepSaturate>(a, stream);
  }
}

hipError_t launch_blur_down(const SepArgs &a, hipStream_t stream) {
  if (a.width <= 0 || a.height <= 0 || a.channels <= 0) return hipSuccess;
  const int out_w = (a.width + 1) / 2, out_h = (a.height + 1) / 2;
  const int per_tile = a.channels == 3 ? Down<3>::kOutW : Down<4>::kOutW;
  const int tx = (out_w * a.channels + per_tile - 1) / per_tile;
  const int ty = (out_h + kDownRows - 1) / kDownRows;
  dim3 grid(tx, ty);
  int tiles_x = 0;
  if (int64_t{tx} * ty <= kMaxLinearTiles) {
    grid = dim3(tx * ty);
    tiles_x = tx;
  }
  switch (a.channels) {
    case 1: blur_down_kernel<1><<<grid, kThreads, 0, stream>>>(a, tiles_x); break;
    case 2: blur_down_kernel<2><<<grid, kThreads, 0, stream>>>(a, tiles_x); break;
    case 3: blur_down_kernel<3><<<grid, kThreads, 0, stream>>>(a, tiles_x); break;
    case 4: blur_down_kernel<4><<<grid, kThreads, 0, stream>>>(a, tiles_x); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_gaussian_arbitrary(const GaussArbArgs &a, hipStream_t stream) {
  if (a.width <= 0 || a.height <= 0 || a.channels <= 0) return hipSuccess;
  if (a.ksize < 3 || a.ksize > 255 || (a.ksize & 1) == 0) return hipErrorInvalidValue;
  const int row_bytes = a.width * a.channels;
  const int vy = min((a.height + kArbRows - 1) / kArbRows, 65535);
  GaussArbVertical v;
  v.a = a;
  // phase[s][m] byte b: tap 4m + b - s (zero outside [0, K))
  for (int s = 0; s < 4; ++s)
    for (int m = 0; m < kArbVDwords; ++m) {
      uint32_t w = 0;
      for (int b = 0; b < 4; ++b) {
        const int i = 4 * m + b - s;
        if (i >= 0 && i < a.ksize) w |= static_cast<uint32_t>(a.taps[i]) << (8 * b);
      }
      v.phase[s * kArbVDwords + m] = w;
    }
  gauss_arb_vertical_kernel<<<dim3((row_bytes + 4 * kThreads - 1) / (4 * kThreads), vy),
                              kThreads, 0, stream>>>(v);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  gauss_arb_horizontal_kernel<<<dim3((row_bytes + kArbRowBytes - 1) / kArbRowBytes,
                                     min(a.height, 65535)),
                                kThreads, 0, stream>>>(a);
  return hipGetLastError();
}

hipError_t launch_median(const MedianArgs &a, hipStream_t stream) {
  if (a.width <= 0 || a.height <= 0 || a.channels <= 0) return hipSuccess;
  if (a.ksize > kMedianNetworkMax) {
    if (a.ksize > 255 || (a.ksize & 1) == 0) return hipErrorInvalidValue;
    const int bx = (a.width * a.channels + kHistLanes - 1) / kHistLanes;
    const int strip = median_hist_strip(bx, a.height, a.ksize, runtime().compute_units);
    const dim3 grid(bx, (a.height + strip - 1) / strip);
    switch (a.channels) {
      case 1: median_hist_kernel<1><<<grid, kHistLanes, 0, stream>>>(a, strip); break;
      case 2: median_hist_kernel<2><<<grid, kHistLanes, 0, stream>>>(a, strip); break;
      case 3: median_hist_kernel<3><<<grid, kHistLanes, 0, stream>>>(a, strip); break;
      case 4: median_hist_kernel<4><<<grid, kHistLanes, 0, stream>>>(a, strip); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  const TileGrid g = tile_grid(a.width, a.height, a.channels, kTileH);
  switch (a.ksize * 8 + a.channels) {
#define KCV_M(K, C) \
  case K * 8 + C: median_kernel<K, C><<<g.grid, kThreads, 0, stream>>>(a, g.tiles_x); break;
    KCV_M(3, 1) KCV_M(3, 2) KCV_M(3, 3) KCV_M(3, 4)
    KCV_M(5, 1) KCV_M(5, 2) KCV_M(5, 3) KCV_M(5, 4)
    KCV_M(7, 1) KCV_M(7, 2) KCV_M(7, 3) KCV_M(7, 4)
    KCV_M(9, 1) KCV_M(9, 2) KCV_M(9, 3) KCV_M(9, 4)
    KCV_M(11, 1) KCV_M(11, 2) KCV_M(11, 3) KCV_M(11, 4)
#undef KCV_M
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace kcv
