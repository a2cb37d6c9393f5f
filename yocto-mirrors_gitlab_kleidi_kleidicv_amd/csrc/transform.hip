// Per-image launches of the geometric gathers (bodies in transform_body.h).

#include "transform_body.h"

namespace kcv {
namespace {

using namespace xform;

// The grid's y dimension is capped at 65535 rows of tiles; blocks stride
// over the rest.
template <int C, int BORDER>
__global__ __launch_bounds__(kThreads) void remap_s16point5_kernel(RemapS16Args a) {
  for (int by = blockIdx.y; by * kBlockPx < a.dst_height; by += gridDim.y)
    remap_s16point5_tile<C, BORDER>(a, blockIdx.x, by);
}

template <int C, int INTERP, int BORDER>
__global__ __launch_bounds__(kThreads) void remap_f32_kernel(RemapF32Args a) {
  for (int by = blockIdx.y; by * kBlockPx < a.dst_height; by += gridDim.y)
    remap_f32_tile<C, INTERP, BORDER>(a, blockIdx.x, by);
}

template <int C, int INTERP, int BORDER>
__global__ __launch_bounds__(kThreads) void warp_kernel(WarpArgs a) {
  for (int by = blockIdx.y; by * kBlockPx < a.dst_height; by += gridDim.y)
    warp_tile<C, INTERP, BORDER>(a, blockIdx.x, by);
}

dim3 tile_grid(int width, int height) {
  const int gy = (height + kBlockPx - 1) / kBlockPx;
  return dim3((width + kBlockPx - 1) / kBlockPx, gy < 65535 ? gy : 65535);
}

#define KCV_FLOAT_DISPATCH(KERNEL, ARGS, C, INTERP, BORDER, GRID, STREAM)      \
  do {                                                                          \
    auto go = [&](auto c_tag, auto i_tag, auto b_tag) {                         \
      KERNEL<decltype(c_tag)::value, decltype(i_tag)::value,                   \
             decltype(b_tag)::value><<<GRID, kThreads, 0, STREAM>>>(ARGS);      \
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
  const dim3 grid = tile_grid(a.dst_width, a.dst_height);
  if (a.channels == 1) {
    if (a.border == kReplicate)
      remap_s16point5_kernel<1, kReplicate><<<grid, kThreads, 0, stream>>>(a);
    else
      remap_s16point5_kernel<1, kConstant><<<grid, kThreads, 0, stream>>>(a);
  } else {
    if (a.border == kReplicate)
      remap_s16point5_kernel<4, kReplicate><<<grid, kThreads, 0, stream>>>(a);
    else
      remap_s16point5_kernel<4, kConstant><<<grid, kThreads, 0, stream>>>(a);
  }
  return hipGetLastError();
}

hipError_t launch_remap_f32(const RemapF32Args &a, hipStream_t stream) {
  if (a.dst_width <= 0 || a.dst_height <= 0) return hipSuccess;
  const dim3 grid = tile_grid(a.dst_width, a.dst_height);
  KCV_FLOAT_DISPATCH(remap_f32_kernel, a, a.channels, a.interpolation, a.border, grid,
                     stream);
  return hipGetLastError();
}

hipError_t launch_warp_perspective(const WarpArgs &a, hipStream_t stream) {
  if (a.dst_width <= 0 || a.dst_height <= 0) return hipSuccess;
  const dim3 grid = tile_grid(a.dst_width, a.dst_height);
  KCV_FLOAT_DISPATCH(warp_kernel, a, a.channels, a.interpolation, a.border, grid, stream);
  return hipGetLastError();
}

}  // namespace kcv
