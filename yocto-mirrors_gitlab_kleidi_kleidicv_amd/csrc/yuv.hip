// Per-image launch of semi-planar YUV to RGB (body in yuv_body.h).

#include "yuv_body.h"

namespace kcv {
namespace {

using namespace yuv;

template <int OC>
__global__ __launch_bounds__(kThreads) void nv12_kernel(YuvArgs a) {
  nv12_rows<OC>(a, blockIdx.x, blockIdx.y, (a.height + 1) / 2, gridDim.y);
}

}  // namespace

hipError_t launch_yuv_sp_to_rgb(const YuvArgs &a, hipStream_t stream) {
  if (a.width <= 0 || a.height <= 0) return hipSuccess;
  const int per_block = kThreads * kBlockW;
  const int pairs = (a.height + 1) / 2;
  const dim3 grid((a.width + per_block - 1) / per_block, min(pairs, 65535));
  if (a.alpha)
    nv12_kernel<4><<<grid, kThreads, 0, stream>>>(a);
  else
    nv12_kernel<3><<<grid, kThreads, 0, stream>>>(a);
  return hipGetLastError();
}

}  // namespace kcv
