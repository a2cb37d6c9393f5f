// Per-image launches of the geometric gathers (bodies in transform_body.h).

#include "transform_body.h"

namespace kcv {
namespace {

using namespace xform;

// Tiles of this workgroup: one XCD-ordered tile of a 1-D grid (neighbouring
// tiles, whose source footprints overlap, then share one XCD's L2), or with
// tiles_x == 0 a 2-D grid whose y dimension is capped at 65535 rows of
// tiles, blocks striding over the rest.  (The kernel argument is passed on
// by value: taking its address would copy it to scratch memory.)
#define KCV_FOR_TILES(TILES_X, DST_HEIGHT, TILE_H, CALL)              \
  do {                                                                \
    if ((TILES_X) > 0) {                                              \
      const int t_ = xcd_tile(blockIdx.x, gridDim.x);                 \
      const int by = t_ / (TILES_X);                                  \
      const int bx = t_ - by * (TILES_X);                             \
      CALL;                                                           \
    } else {                                                          \
      const int bx = blockIdx.x;                                      \
      for (int by = blockIdx.y; by * (TILE_H) < (DST_HEIGHT); by += gridDim.y) CALL; \
    }                                                                 \
  } while (0)

template <int C, int BORDER>
__global__ __launch_bounds__(kThreads) void remap_s16point5_kernel(RemapS16Args a,
                                                                   int tiles_x) {
  KCV_FOR_TILES(tiles_x, a.dst_height, gather_tile_h(C), (remap_s16point5_tile<C, BORDER>(a, bx, by)));
}

template <int C, int INTERP, int BORDER>
__global__ __launch_bounds__(kThreads, float_gather_blocks_per_cu(C)) void remap_f32_kernel(RemapF32Args a,
                                                                              int tiles_x) {
  KCV_FOR_TILES(tiles_x, a.dst_height, gather_tile_h(C), (remap_f32_tile<C, INTERP, BORDER>(a, bx, by)));
}

template <int C, int INTERP, int BORDER>
__global__ __launch_bounds__(kThreads, float_gather_blocks_per_cu(C)) void warp_kernel(WarpArgs a, int tiles_x) {
  KCV_FOR_TILES(tiles_x, a.dst_height, gather_tile_h(C), (warp_tile<C, INTERP, BORDER>(a, bx, by)));
}

struct TileGrid {
  dim3 grid;
  int tiles_x;
};

TileGrid tile_grid(int width, int height, int channels) {
  const int tile_w = gather_tile_w(channels), tile_h = gather_tile_h(channels);
  const int tx = (width + tile_w - 1) / tile_w;
  const int ty = (height + tile_h - 1) / tile_h;
  if (int64_t{tx} * ty <= kMaxLinearTiles) return {dim3(tx * ty), tx};
  return {dim3(tx, ty < 65535 ? ty : 65535), 0};
}

#define KCV_FLOAT_DISPATCH(KERNEL, ARGS, TX, C, INTERP, BORDER, GRID, STREAM)  \
  do {                                                                          \
    auto go = [&](auto c_tag, auto i_tag, auto b_tag) {                         \
      KERNEL<decltype(c_tag)::value, decltype(i_tag)::value,                   \
             decltype(b_tag)::value><<<GRID, kThreads, 0, STREAM>>>(ARGS, TX);  \
    };                                                                          \
    auto with_b = [&](auto c_tag, auto i_tag) {                                 \
      if ((BORDER) == kReplicate)                                               \
        go(c_tag, i_tag, std::integral_constant<int, kReplicate>{});           \
      else                                                                      \
        go(c_tag, i_tag, std::integral_constant<int, kConstant>{});            \
    };                                                                          \
    auto with_i = [&](auto c_tag) {                                             \
      if ((INTERP) == 1)                                                        \
        with_b(c_tag, std::integral_constant<int, 1>{});                       \
      else                                                                      \
        with_b(c_tag, std::integral_constant<int, 0>{});                       \
    };                                                                          \
    switch (C) {                                                                \
      case 1: with_i(std::integral_constant<int, 1>{}); break;                 \
      case 2: with_i(std::integral_constant<int, 2>{}); break;                 \
      case 3: with_i(std::integral_constant<int, 3>{}); break;                 \
      default: with_i(std::integral_constant<int, 4>{}); break;                \
    }                                                                           \
  } while (0)

}  // namespace

hipError_t launch_remap_s16point5(const RemapS16Args &a, hipStream_t stream) {
  if (a.dst_width <= 0 || a.dst_height <= 0) return hipSuccess;
  const TileGrid g = tile_grid(a.dst_width, a.dst_height, a.channels);
  if (a.channels == 1) {
    if (a.border == kReplicate)
      remap_s16point5_kernel<1, kReplicate><<<g.grid, kThreads, 0, stream>>>(a, g.tiles_x);
    else
      remap_s16point5_kernel<1, kConstant><<<g.grid, kThreads, 0, stream>>>(a, g.tiles_x);
  } else {
    if (a.border == kReplicate)
      remap_s16point5_kernel<4, kReplicate><<<g.grid, kThreads, 0, stream>>>(a, g.tiles_x);
    else
      remap_s16point5_kernel<4, kConstant><<<g.grid, kThreads, 0, stream>>>(a, g.tiles_x);
  }
  return hipGetLastError();
}

hipError_t launch_remap_f32(const RemapF32Args &a, hipStream_t stream) {
  if (a.dst_width <= 0 || a.dst_height <= 0) return hipSuccess;
  const TileGrid g = tile_grid(a.dst_width, a.dst_height, a.channels);
  KCV_FLOAT_DISPATCH(remap_f32_kernel, a, g.tiles_x, a.channels, a.interpolation, a.border,
                     g.grid, stream);
  return hipGetLastError();
}

hipError_t launch_warp_perspective(const WarpArgs &a, hipStream_t stream) {
  if (a.dst_width <= 0 || a.dst_height <= 0) return hipSuccess;
  const TileGrid g = tile_grid(a.dst_width, a.dst_height, a.channels);
  KCV_FLOAT_DISPATCH(warp_kernel, a, g.tiles_x, a.channels, a.interpolation, a.border,
                     g.grid, stream);
  return hipGetLastError();
}

namespace {

// recip_refined against the IEEE quotient for the 2^24 float bit patterns
// starting at `base`; lanes outside recip_refined_ok are skipped.
__global__ __launch_bounds__(kThreads) void reciprocal_check_kernel(uint32_t base,
                                                                    unsigned long long *bad) {
  const uint32_t bits = base + blockIdx.x * kThreads + threadIdx.x;
  const float w = __builtin_bit_cast(float, bits);
  if (!recip_refined_ok(w)) return;
  const float q = 1.0f / w;
  if (__builtin_bit_cast(uint32_t, recip_refined(w)) != __builtin_bit_cast(uint32_t, q))
    atomicAdd(bad, 1ull);
}

// round_half_away_u8 against truncation plus a half test for the float bit
// patterns [base, min(base + 2^24, limit)].
__global__ __launch_bounds__(kThreads) void rounding_check_kernel(uint32_t base, uint32_t limit,
                                                                  unsigned long long *bad) {
  const uint32_t bits = base + blockIdx.x * kThreads + threadIdx.x;
  if (bits > limit) return;
  const float x = __builtin_bit_cast(float, bits);
  const float t = truncf(x);
  const uint32_t want = static_cast<uint32_t>(t) + (x - t >= 0.5f ? 1u : 0u);
  if (round_half_away_u8(x) != want) atomicAdd(bad, 1ull);
}

}  // namespace
}  // namespace kcv

// Self-check of the lerp rounding used by remap f32 and warp perspective
// (not part of the KleidiCV API): counts the floats x in [0, 256] where the
// hardware conversion differs from round-half-away.  Synchronous; returns 0
// on success.
extern "C" int kcv_selfcheck_rounding(unsigned long long *mismatches) {
  unsigned long long *bad = nullptr;
  if (hipMalloc(&bad, sizeof(*bad)) != hipSuccess) return 1;
  hipError_t e = hipMemset(bad, 0, sizeof(*bad));
  constexpr uint32_t kChunk = 1u << 24, kLimit = 0x43800000u;  // 256.0f
  for (uint64_t base = 0; e == hipSuccess && base <= kLimit; base += kChunk) {
    kcv::rounding_check_kernel<<<kChunk / kcv::xform::kThreads, kcv::xform::kThreads>>>(
        static_cast<uint32_t>(base), kLimit, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(mismatches, bad, sizeof(*bad), hipMemcpyDeviceToHost);
  const hipError_t freed = hipFree(bad);
  return e == hipSuccess && freed == hipSuccess ? 0 : 1;
}

// Self-check of the reciprocal used by warp perspective (not part of the
// KleidiCV API): counts the floats w with |w| in [2^-125, 2^125] whose
// refined hardware reciprocal differs from 1.0f / w.  Synchronous; returns
// 0 on success.
extern "C" int kcv_selfcheck_reciprocal(unsigned long long *mismatches) {
  unsigned long long *bad = nullptr;
  if (hipMalloc(&bad, sizeof(*bad)) != hipSuccess) return 1;
  hipError_t e = hipMemset(bad, 0, sizeof(*bad));
  constexpr uint32_t kChunk = 1u << 24;
  for (uint64_t base = 0; e == hipSuccess && base < (uint64_t{1} << 32); base += kChunk) {
    kcv::reciprocal_check_kernel<<<kChunk / kcv::xform::kThreads, kcv::xform::kThreads>>>(
        static_cast<uint32_t>(base), bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(mismatches, bad, sizeof(*bad), hipMemcpyDeviceToHost);
  const hipError_t freed = hipFree(bad);
  return e == hipSuccess && freed == hipSuccess ? 0 : 1;
}
