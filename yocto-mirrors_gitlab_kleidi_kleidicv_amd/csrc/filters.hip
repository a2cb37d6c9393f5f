// Per-image launches of the tiled 2-D filters (bodies in filters_body.h).

#include "filters_body.h"

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
  __shared__ __align__(16) uint8_t smem[separable_smem_bytes<K>()];
  int bx, by;
  tile_of(tiles_x, bx, by);
  separable_tile<MODE, K, C>(a, bx, by, smem);
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
__global__ __launch_bounds__(kThreads) void median_kernel(MedianArgs a, int tiles_x) {
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

// Grid and tiles_x argument for an image: a 1-D XCD-ordered grid unless the
// tile count is very large.
struct TileGrid {
  dim3 grid;
  int tiles_x;
};

TileGrid tile_grid(int width, int height, int channels) {
  const int tx = (width * channels + kTileW - 1) / kTileW;
  const int ty = (height + kTileH - 1) / kTileH;
  if (int64_t{tx} * ty <= kMaxLinearTiles) return {dim3(tx * ty), tx};
  return {dim3(tx, ty), 0};
}

template <int MODE>
hipError_t launch_mode(const SepArgs &a, hipStream_t s) {
  const TileGrid g = tile_grid(a.width, a.height, a.channels);
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

hipError_t launch_median(const MedianArgs &a, hipStream_t stream) {
  if (a.width <= 0 || a.height <= 0 || a.channels <= 0) return hipSuccess;
  const TileGrid g = tile_grid(a.width, a.height, a.channels);
  switch (a.ksize * 8 + a.channels) {
#define KCV_M(K, C) \
  case K * 8 + C: median_kernel<K, C><<<g.grid, kThreads, 0, stream>>>(a, g.tiles_x); break;
    KCV_M(3, 1) KCV_M(3, 2) KCV_M(3, 3) KCV_M(3, 4)
    KCV_M(5, 1) KCV_M(5, 2) KCV_M(5, 3) KCV_M(5, 4)
    KCV_M(7, 1) KCV_M(7, 2) KCV_M(7, 3) KCV_M(7, 4)
#undef KCV_M
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace kcv
