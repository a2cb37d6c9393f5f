// Per-image launches of the tiled 2-D filters (bodies in filters_body.h).

#include "filters_body.h"

namespace kcv {
namespace {

using namespace filt;

template <int MODE, int K>
__global__ __launch_bounds__(kThreads) void separable_kernel(SepArgs a) {
  __shared__ __align__(16) uint8_t smem[separable_smem_bytes<K>()];
  separable_tile<MODE, K>(a, blockIdx.x, blockIdx.y, smem);
}

template <int K>
__global__ __launch_bounds__(kThreads) void median_kernel(MedianArgs a) {
  __shared__ __align__(16) uint8_t smem[median_smem_bytes<K>()];
  median_tile<K>(a, blockIdx.x, blockIdx.y, smem);
}

dim3 tile_grid(int width, int height, int channels) {
  return dim3((width * channels + kTileW - 1) / kTileW,
              (height + kTileH - 1) / kTileH);
}

template <int MODE>
hipError_t launch_mode(const SepArgs &a, hipStream_t s) {
  const dim3 grid = tile_grid(a.width, a.height, a.channels);
  switch (a.ksize) {
    case 3: separable_kernel<MODE, 3><<<grid, kThreads, 0, s>>>(a); break;
    case 5: separable_kernel<MODE, 5><<<grid, kThreads, 0, s>>>(a); break;
    case 7: separable_kernel<MODE, 7><<<grid, kThreads, 0, s>>>(a); break;
    case 9:
      if constexpr (MODE != kSepSaturate) {
        separable_kernel<MODE, 9><<<grid, kThreads, 0, s>>>(a);
        break;
      }
      return hipErrorInvalidValue;
    case 15:
      if constexpr (MODE == kSepSigma) {
        separable_kernel<MODE, 15><<<grid, kThreads, 0, s>>>(a);
        break;
      }
      return hipErrorInvalidValue;
    case 21:
      if constexpr (MODE == kSepSigma) {
        separable_kernel<MODE, 21><<<grid, kThreads, 0, s>>>(a);
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

hipError_t launch_median(const MedianArgs &a, hipStream_t stream) {
  if (a.width <= 0 || a.height <= 0 || a.channels <= 0) return hipSuccess;
  const dim3 grid = tile_grid(a.width, a.height, a.channels);
  switch (a.ksize) {
    case 3: median_kernel<3><<<grid, kThreads, 0, stream>>>(a); break;
    case 5: median_kernel<5><<<grid, kThreads, 0, stream>>>(a); break;
    case 7: median_kernel<7><<<grid, kThreads, 0, stream>>>(a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace kcv
