// Per-image launches of YUV to RGB: semi-planar (NV12 / NV21), planar 4:2:0,
// interleaved 4:2:2 and 4:4:4 (bodies in yuv_body.h).

#include "yuv_body.h"

namespace kcv {
namespace {

using namespace yuv;

template <int OC>
__global__ __launch_bounds__(kThreads) void nv12_kernel(YuvArgs a) {
  nv12_rows<OC>(a, blockIdx.x, blockIdx.y, (a.height + 1) / 2, gridDim.y);
}

template <int OC>
__global__ __launch_bounds__(kThreads) void yuv420p_kernel(YuvXArgs a) {
  yuv420p_rows<OC>(a, blockIdx.x, blockIdx.y, gridDim.y);
}

template <int OC>
__global__ __launch_bounds__(kThreads) void yuv422_kernel(YuvXArgs a) {
  yuv422_rows<OC>(a, blockIdx.x, blockIdx.y, gridDim.y);
}

template <int OC>
__global__ __launch_bounds__(kThreads) void yuv444_kernel(YuvXArgs a) {
  yuv444_rows<OC>(a, blockIdx.x, blockIdx.y, gridDim.y);
}

template <int SCN>
__global__ __launch_bounds__(kThreads) void rgb_to_yuv420_kernel(RgbYuvArgs a) {
  rgb_to_yuv420_rows<SCN>(a, blockIdx.x, blockIdx.y, gridDim.y);
}

template <int SCN>
__global__ __launch_bounds__(kThreads) void rgb_to_yuv422_kernel(RgbYuvArgs a) {
  rgb_to_yuv422_rows<SCN>(a, blockIdx.x, blockIdx.y, gridDim.y);
}

template <int SCN>
__global__ __launch_bounds__(kThreads) void rgb_to_yuv444_kernel(RgbYuvArgs a) {
  rgb_to_yuv444_rows<SCN>(a, blockIdx.x, blockIdx.y, gridDim.y);
}

}  // namespace

hipError_t launch_rgb_to_yuv(const RgbYuvArgs &a, hipStream_t stream) {
  if (a.width <= 0 || a.height <= 0) return hipSuccess;
  const int per_block = kThreads * kBlockW;
  const bool is420 = a.layout == kYuv420P || a.layout == kRgb420SP;
  const int rows = is420 ? (a.height + 1) / 2 : a.height;
  const dim3 grid((a.width + per_block - 1) / per_block, min(rows, 65535));
  const bool four = a.scn == 4;
  if (is420) {
    if (four) rgb_to_yuv420_kernel<4><<<grid, kThreads, 0, stream>>>(a);
    else rgb_to_yuv420_kernel<3><<<grid, kThreads, 0, stream>>>(a);
  } else if (a.layout == kYuv422) {
    if (four) rgb_to_yuv422_kernel<4><<<grid, kThreads, 0, stream>>>(a);
    else rgb_to_yuv422_kernel<3><<<grid, kThreads, 0, stream>>>(a);
  } else if (a.layout == kYuv444) {
    if (four) rgb_to_yuv444_kernel<4><<<grid, kThreads, 0, stream>>>(a);
    else rgb_to_yuv444_kernel<3><<<grid, kThreads, 0, stream>>>(a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_yuv_to_rgb(const YuvXArgs &a, hipStream_t stream) {
  if (a.width <= 0 || a.height <= 0) return hipSuccess;
  const int per_block = kThreads * kBlockW;
  const int rows = a.layout == kYuv420P ? (a.height + 1) / 2 : a.height;
  const dim3 grid((a.width + per_block - 1) / per_block, min(rows, 65535));
  switch (a.layout * 2 + (a.alpha ? 1 : 0)) {
    case kYuv420P * 2: yuv420p_kernel<3><<<grid, kThreads, 0, stream>>>(a); break;
    case kYuv420P * 2 + 1: yuv420p_kernel<4><<<grid, kThreads, 0, stream>>>(a); break;
    case kYuv422 * 2: yuv422_kernel<3><<<grid, kThreads, 0, stream>>>(a); break;
    case kYuv422 * 2 + 1: yuv422_kernel<4><<<grid, kThreads, 0, stream>>>(a); break;
    case kYuv444 * 2: yuv444_kernel<3><<<grid, kThreads, 0, stream>>>(a); break;
    case kYuv444 * 2 + 1: yuv444_kernel<4><<<grid, kThreads, 0, stream>>>(a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

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
