// Geometric gathers for gfx950: remap with s16 + 5/5-bit fractional maps,
// remap with float maps and warp-perspective.
//
// Each thread owns 4 consecutive destination pixels of one row; map reads
// and destination writes are coalesced across the wavefront, source reads
// are gathers that mostly hit L1/L2 for smooth maps.
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

#include "kcv_internal.h"

namespace kcv {
namespace {

constexpr int kThreads = 256;
constexpr int kPixelsPerThread = 4;

// ---------------------------------------------------------------------------
// remap s16point5
__device__ __forceinline__ uint32_t bilinear5(uint32_t a, uint32_t b,
                                              uint32_t c, uint32_t d,
                                              uint32_t fx, uint32_t fy) {
  const uint32_t line0 = (32 - fx) * a + fx * b;
  const uint32_t line1 = (32 - fx) * c + fx * d;
  return (line0 * (32 - fy) + line1 * fy + 512) >> 10;
}

template <int C, int BORDER>
__global__ __launch_bounds__(kThreads) void remap_s16point5_kernel(
    RemapS16Args a) {
  const int x0 = (blockIdx.x * kThreads + threadIdx.x) * kPixelsPerThread;
  if (x0 >= a.dst_width) return;
  for (int y = blockIdx.y; y < a.dst_height; y += gridDim.y) {
  const int16_t *mrow = reinterpret_cast<const int16_t *>(
      reinterpret_cast<const uint8_t *>(a.mapxy) + y * a.mapxy_stride);
  const uint16_t *frow = reinterpret_cast<const uint16_t *>(
      reinterpret_cast<const uint8_t *>(a.mapfrac) + y * a.mapfrac_stride);
  uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride;
  const int xmax = a.src_width - 1, ymax = a.src_height - 1;
#pragma unroll
  for (int k = 0; k < kPixelsPerThread; ++k) {
    const int x = x0 + k;
    if (x >= a.dst_width) break;
    const int mx = mrow[2 * x], my = mrow[2 * x + 1];
    const uint32_t f = frow[x];
    uint32_t fx = f & 31u, fy = (f >> 5) & 31u;
    if (BORDER == kReplicate) {
      // RemapS16Point5Replicate (:16-137) / get_coordinates_replicate
      // (:621-640): fractions of negative coordinates are dropped, taps
      // clamp to the image.
      if (mx < 0) fx = 0;
      if (my < 0) fy = 0;
      const int cx0 = min(max(mx, 0), xmax), cy0 = min(max(my, 0), ymax);
      const int cx1 = cx0 + (cx0 < xmax ? 1 : 0);
      const int cy1 = cy0 + (cy0 < ymax ? 1 : 0);
      const uint8_t *r0 = a.src + static_cast<size_t>(cy0) * a.src_stride;
      const uint8_t *r1 = a.src + static_cast<size_t>(cy1) * a.src_stride;
#pragma unroll
      for (int c = 0; c < C; ++c)
        out[x * C + c] = static_cast<uint8_t>(bilinear5(
            r0[cx0 * C + c], r0[cx1 * C + c], r1[cx0 * C + c], r1[cx1 * C + c],
            fx, fy));
    } else {
      // RemapS16Point5ConstantBorder (:321-425) / get_coordinates_constant
      // (:860-881): each corner tests its own range with 16-bit unsigned
      // compares and takes the border value when outside.
      const uint32_t ux0 = static_cast<uint16_t>(mx), uy0 = static_cast<uint16_t>(my);
      const uint32_t ux1 = static_cast<uint16_t>(ux0 + 1);
      const uint32_t uy1 = static_cast<uint16_t>(uy0 + 1);
      const bool x0in = ux0 < static_cast<uint32_t>(a.src_width);
      const bool x1in = ux1 < static_cast<uint32_t>(a.src_width);
      const bool y0in = uy0 < static_cast<uint32_t>(a.src_height);
      const bool y1in = uy1 < static_cast<uint32_t>(a.src_height);
      const uint8_t *r0 = a.src + static_cast<size_t>(y0in ? uy0 : 0) * a.src_stride;
      const uint8_t *r1 = a.src + static_cast<size_t>(y1in ? uy1 : 0) * a.src_stride;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const uint32_t bv = a.border_value[C == 4 ? c : 0];
        const uint32_t pa = (x0in && y0in) ? r0[ux0 * C + c] : bv;
        const uint32_t pb = (x1in && y0in) ? r0[ux1 * C + c] : bv;
        const uint32_t pc = (x0in && y1in) ? r1[ux0 * C + c] : bv;
        const uint32_t pd = (x1in && y1in) ? r1[ux1 * C + c] : bv;
        out[x * C + c] = static_cast<uint8_t>(min(bilinear5(pa, pb, pc, pd, fx, fy), 255u));
      }
    }
  }
  }
}

// ---------------------------------------------------------------------------
// Float gathers (transform_neon.h:189-491).
__device__ __forceinline__ uint32_t cvtm_u32(float f) {
  if (!(f == f)) return 0;
  const float r = floorf(f);
  if (r <= 0.0f) return 0;
  if (r >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(r);
}

__device__ __forceinline__ int32_t cvtm_s32(float f) {
  if (!(f == f)) return 0;
  const float r = floorf(f);
  if (r >= 2147483648.0f) return 0x7FFFFFFF;
  if (r < -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(r);
}

__device__ __forceinline__ uint32_t cvta_u32(float f) {
  if (!(f == f)) return 0;
  const float r = roundf(f);  // ties away from zero
  if (r <= 0.0f) return 0;
  if (r >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(r);
}

__device__ __forceinline__ int32_t cvta_s32(float f) {
  if (!(f == f)) return 0;
  const float r = roundf(f);
  if (r >= 2147483648.0f) return 0x7FFFFFFF;
  if (r < -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(r);
}

struct Src {
  const uint8_t *p;
  size_t stride;
  uint32_t xmax, ymax;
};

// load_quad_pixels_{replicate,constant} + lerp_2d, all channels of one pixel.
template <int C, int BORDER>
__device__ __forceinline__ void sample_linear(const Src &s, float xf, float yf,
                                              const uint8_t *bv, uint8_t *out) {
  const float xfrac = xf - floorf(xf), yfrac = yf - floorf(yf);
  uint32_t x0, y0, x1, y1;
  bool ina = true, inb = true, inc = true, ind = true;
  if (BORDER == kReplicate) {
    x0 = min(cvtm_u32(xf), s.xmax);
    y0 = min(cvtm_u32(yf), s.ymax);
    x1 = x0 + ((xf >= 0.0f && x0 < s.xmax) ? 1u : 0u);
    y1 = y0 + ((yf >= 0.0f && y0 < s.ymax) ? 1u : 0u);
  } else {
    x0 = static_cast<uint32_t>(cvtm_s32(xf));
    y0 = static_cast<uint32_t>(cvtm_s32(yf));
    x1 = x0 + 1u;
    y1 = y0 + 1u;
    const bool x0in = x0 <= s.xmax, x1in = x1 <= s.xmax;
    const bool y0in = y0 <= s.ymax, y1in = y1 <= s.ymax;
    ina = x0in && y0in;
    inb = x1in && y0in;
    inc = x0in && y1in;
    ind = x1in && y1in;
  }
  const uint8_t *r0 = s.p + static_cast<size_t>(BORDER == kReplicate || y0 <= s.ymax ? y0 : 0) * s.stride;
  const uint8_t *r1 = s.p + static_cast<size_t>(BORDER == kReplicate || y1 <= s.ymax ? y1 : 0) * s.stride;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const float pa = static_cast<float>(ina ? r0[static_cast<size_t>(x0) * C + c] : bv[c]);
    const float pb = static_cast<float>(inb ? r0[static_cast<size_t>(x1) * C + c] : bv[c]);
    const float pc = static_cast<float>(inc ? r1[static_cast<size_t>(x0) * C + c] : bv[c]);
    const float pd = static_cast<float>(ind ? r1[static_cast<size_t>(x1) * C + c] : bv[c]);
    const float line0 = fmaf(pb - pa, xfrac, pa);
    const float line1 = fmaf(pd - pc, xfrac, pc);
    const float res = fmaf(line1 - line0, yfrac, line0);
    out[c] = static_cast<uint8_t>(cvta_u32(res));
  }
}

template <int C, int BORDER>
__device__ __forceinline__ void sample_nearest(const Src &s, float xf, float yf,
                                               const uint8_t *bv, uint8_t *out) {
  if (BORDER == kReplicate) {
    const uint32_t x = min(cvta_u32(xf), s.xmax);
    const uint32_t y = min(cvta_u32(yf), s.ymax);
    const uint8_t *p = s.p + static_cast<size_t>(y) * s.stride + static_cast<size_t>(x) * C;
#pragma unroll
    for (int c = 0; c < C; ++c) out[c] = p[c];
  } else {
    const uint32_t x = static_cast<uint32_t>(cvta_s32(xf));
    const uint32_t y = static_cast<uint32_t>(cvta_s32(yf));
    const bool in = x <= s.xmax && y <= s.ymax;
    const uint8_t *p = s.p + (in ? static_cast<size_t>(y) * s.stride + static_cast<size_t>(x) * C : 0);
#pragma unroll
    for (int c = 0; c < C; ++c) out[c] = in ? p[c] : bv[c];
  }
}

template <int C, int INTERP, int BORDER>
__global__ __launch_bounds__(kThreads) void remap_f32_kernel(RemapF32Args a) {
  const int x0 = (blockIdx.x * kThreads + threadIdx.x) * kPixelsPerThread;
  if (x0 >= a.dst_width) return;
  for (int y = blockIdx.y; y < a.dst_height; y += gridDim.y) {
  const float *mx = reinterpret_cast<const float *>(
      reinterpret_cast<const uint8_t *>(a.mapx) + y * a.mapx_stride);
  const float *my = reinterpret_cast<const float *>(
      reinterpret_cast<const uint8_t *>(a.mapy) + y * a.mapy_stride);
  uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride;
  const Src s{a.src, a.src_stride, static_cast<uint32_t>(a.src_width - 1),
              static_cast<uint32_t>(a.src_height - 1)};
#pragma unroll
  for (int k = 0; k < kPixelsPerThread; ++k) {
    const int x = x0 + k;
    if (x >= a.dst_width) break;
    if (INTERP == 1)
      sample_linear<C, BORDER>(s, mx[x], my[x], a.border_value, out + x * C);
    else
      sample_nearest<C, BORDER>(s, mx[x], my[x], a.border_value, out + x * C);
  }
  }
}

template <int C, int INTERP, int BORDER>
__global__ __launch_bounds__(kThreads) void warp_kernel(WarpArgs a) {
  const int x0 = (blockIdx.x * kThreads + threadIdx.x) * kPixelsPerThread;
  if (x0 >= a.dst_width) return;
  for (int y = blockIdx.y; y < a.dst_height; y += gridDim.y) {
  // warp_perspective_neon.cpp:98-102: row constants (fused).
  const float fy = static_cast<float>(y);
  const float tx0 = fmaf(a.t[1], fy, a.t[2]);
  const float ty0 = fmaf(a.t[4], fy, a.t[5]);
  const float tw0 = fmaf(a.t[7], fy, a.t[8]);
  uint8_t *out = a.dst + static_cast<size_t>(y) * a.dst_stride;
  const Src s{a.src, a.src_stride, static_cast<uint32_t>(a.src_width - 1),
              static_cast<uint32_t>(a.src_height - 1)};
#pragma unroll
  for (int k = 0; k < kPixelsPerThread; ++k) {
    const int x = x0 + k;
    if (x >= a.dst_width) break;
    // calculate_coordinates (:55-76)
    const float fx = static_cast<float>(static_cast<uint32_t>(x));
    const float tx = fmaf(fx, a.t[0], tx0);
    const float ty = fmaf(fx, a.t[3], ty0);
    const float tw = fmaf(fx, a.t[6], tw0);
    const float iw = (a.check_zero && tw == 0.0f) ? 0.0f : 1.0f / tw;
    const float xf = tx * iw, yf = ty * iw;
    if (INTERP == 1)
      sample_linear<C, BORDER>(s, xf, yf, a.border_value, out + x * C);
    else
      sample_nearest<C, BORDER>(s, xf, yf, a.border_value, out + x * C);
  }
  }
}

dim3 row_grid(int width, int height) {
  const int per_block = kThreads * kPixelsPerThread;
  return dim3((width + per_block - 1) / per_block, min(height, 65535));
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
  const dim3 grid = row_grid(a.dst_width, a.dst_height);
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
  const dim3 grid = row_grid(a.dst_width, a.dst_height);
  KCV_FLOAT_DISPATCH(remap_f32_kernel, a, a.channels, a.interpolation,
                     a.border, grid, stream);
  return hipGetLastError();
}

hipError_t launch_warp_perspective(const WarpArgs &a, hipStream_t stream) {
  if (a.dst_width <= 0 || a.dst_height <= 0) return hipSuccess;
  const dim3 grid = row_grid(a.dst_width, a.dst_height);
  KCV_FLOAT_DISPATCH(warp_kernel, a, a.channels, a.interpolation, a.border,
                     grid, stream);
  return hipGetLastError();
}

}  // namespace kcv
