// Argument validation and kernel-argument preparation (declared in
// prepare.h).  Validation order per entry point follows the reference's API
// files, cited at each function; the first failing check decides the code.

#include "prepare.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>

namespace kcv {

Runtime &runtime() {
  static Runtime rt;
  static std::once_flag once;
  std::call_once(once, [] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return;
    if (hipGetDevice(&rt.device) != hipSuccess) return;
    if (hipDeviceGetAttribute(&rt.compute_units, hipDeviceAttributeMultiprocessorCount,
                              rt.device) != hipSuccess)
      return;
    if (hipStreamCreateWithFlags(&rt.stream, hipStreamNonBlocking) != hipSuccess) return;
    // Per-call scratch (resize tables, staged host buffers) is stream-ordered
    // (hipMallocAsync / hipFreeAsync).  With the default pool's release
    // threshold of 0 every free returns the memory to the driver and the
    // next allocation maps it again, which waits for the device: about
    // 20 us per small resize call.  Keeping freed blocks in the pool makes
    // the pair a few microseconds; the pool holds at most the peak of
    // outstanding scratch.
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, rt.device) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    (void)hipGetLastError();
    rt.ok = true;
  });
  return rt;
}

kleidicv_error_t from_hip(hipError_t e) {
  if (e == hipSuccess) return KLEIDICV_OK;
  if (std::getenv("KLEIDICV_HIP_VERBOSE"))
    std::fprintf(stderr, "kleidicv_hip: %s\n", hipGetErrorString(e));
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return KLEIDICV_ERROR_ALLOCATION;
  return KLEIDICV_ERROR_NOT_IMPLEMENTED;
}

// Pointers that HIP knows as device or managed memory; anything else is
// host memory.
bool is_device_pointer(const void *p) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

namespace {
// Device allocations still to succeed before the forced failure; negative
// when no failure is armed.
std::atomic<long> g_alloc_fail_after{-1};
}  // namespace

hipError_t device_alloc(void **p, size_t bytes, hipStream_t stream, bool stream_ordered) {
  long n = g_alloc_fail_after.load();
  while (n >= 0 && !g_alloc_fail_after.compare_exchange_weak(n, n - 1)) {
  }
  if (n == 0) {
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  return stream_ordered ? hipMallocAsync(p, bytes, stream) : hipMalloc(p, bytes);
}

namespace {

// Checks shared by the entry points (include/kleidicv/utils.h:374-429).
template <typename T>
kleidicv_error_t check_ptr_stride(const T *p, size_t stride, size_t height) {
  if (p == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  if (alignof(T) > 1 && height > 1 && (stride % alignof(T)) != 0)
    return KLEIDICV_ERROR_ALIGNMENT;
  return KLEIDICV_OK;
}

kleidicv_error_t check_image_size(size_t width, size_t height) {
  size_t n = 0;
  if (__builtin_mul_overflow(width, height, &n)) return KLEIDICV_ERROR_RANGE;
  if (n > KLEIDICV_MAX_IMAGE_PIXELS) return KLEIDICV_ERROR_RANGE;
  return KLEIDICV_OK;
}

bool is_filter_border(kleidicv_border_type_t b) {
  return b == KLEIDICV_BORDER_TYPE_REPLICATE || b == KLEIDICV_BORDER_TYPE_REFLECT ||
         b == KLEIDICV_BORDER_TYPE_WRAP || b == KLEIDICV_BORDER_TYPE_REVERSE;
}

// Small parameter arrays may live in host or device memory.
kleidicv_error_t read_small(void *dst, const void *src, size_t bytes) {
  if (is_device_pointer(src)) return from_hip(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
  std::memcpy(dst, src, bytes);
  return KLEIDICV_OK;
}

// include/kleidicv/filters/gaussian_blur.h:38-82.
bool gaussian_is_implemented(size_t w, size_t h, size_t kw, size_t kh, float sx, float sy,
                             size_t c, kleidicv_border_type_t b) {
  if (kw != kh || kw < 3 || kw > 255 || (kw & 1) != 1 || sx != sy) return false;
  if (w < kw - 1 || h < kw - 1 || c > KLEIDICV_MAXIMUM_CHANNEL_COUNT) return false;
  if (kw > 9 && kw != 15 && kw != 21) {
    if (b != KLEIDICV_BORDER_TYPE_REPLICATE) return false;
    const size_t margin = kw / 2;
    const size_t aligned_margin = (margin + 7) / 8 * 8;
    if (w < aligned_margin + margin) return false;
  }
  return true;
}

// Host half-kernel for sigma != 0 (include/kleidicv/filters/sigma.h:27-79).
// The reference computes the default sigma `ksize * 0.15 + 0.35` and
// `hk * multiplier - error` as single expressions that clang contracts on
// AArch64, hence the explicit fused multiply-adds.
bool gaussian_half_kernel(uint8_t *half, size_t half_size, float sigma) {
  const size_t mid = half_size - 1;
  const size_t ksize = mid * 2 + 1;
  if (sigma == 0.0f) sigma = std::fmaf(static_cast<float>(ksize), 0.15f, 0.35f);
  float vals[128];
  const float coefficient = 1 / -(2 * sigma * sigma);
  float sum = 0.0f;
  for (size_t i = 0, j = mid; i < mid; i++, j--) {
    vals[i] = std::exp(static_cast<float>(j) * static_cast<float>(j) * coefficient);
    sum += vals[i];
  }
  const float multiplier = 256 / std::fmaf(sum, 2.0f, 1.0f);
  float error = 0.0f;
  for (size_t i = 0; i < mid; i++) {
    const float value = std::fmaf(vals[i], multiplier, -error);
    const float rounded = std::round(value);
    half[i] = static_cast<uint8_t>(rounded);
    error = rounded - value;
  }
  const uint16_t mid_value = static_cast<uint16_t>(std::round(multiplier - error));
  if (mid_value == 256) return false;
  half[mid] = static_cast<uint8_t>(mid_value);
  return true;
}

// include/kleidicv/resize/resize_linear.h:16-65.
bool resize_is_implemented(size_t sw, size_t sh, size_t dw, size_t dh, size_t c) {
  const size_t kMax = KLEIDICV_RESIZE_MAX_WIDTH_OR_HEIGHT;
  if (sw > kMax || sh > kMax || dw > kMax || dh > kMax || c == 0) return false;
  if (sw == 0 || sh == 0) return true;
  if (sw == 2 * dw && sh == 2 * dh && c <= 4) return true;
  if (c == 1 && ((sw * 2 == dw && sh * 2 == dh) || (sw * 4 == dw && sh * 4 == dh)))
    return true;
  if (c <= 4 && dw >= sw) return sw * c >= 16;
  if (c <= 4 && dw * 3 >= sw && dw < sw) {
    if (dw * c < 8) return false;
    if (dw * 2 >= sw && c != 3) return sw * c >= 16;
    return sw * c >= 32;
  }
  return false;
}

}  // namespace

// ---------------------------------------------------------------------------
// Gaussian blur.  src/filters/gaussian_blur_api.cpp:21-53,
// include/kleidicv/filters/gaussian_blur.h:38-95.
kleidicv_error_t prepare_gaussian(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                  size_t dst_stride, size_t width, size_t height,
                                  size_t channels, size_t kernel_width,
                                  size_t kernel_height, float sigma_x, float sigma_y,
                                  kleidicv_border_type_t border_type, SepArgs *a,
                                  Action *act, GaussArbArgs *arb) {
  *act = Action::kNone;
  if (!is_filter_border(border_type)) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (!gaussian_is_implemented(width, height, kernel_width, kernel_height, sigma_x,
                               sigma_y, channels, border_type))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  const bool arbitrary = !(kernel_width <= 9 || kernel_width == 15 || kernel_width == 21);
  // the batch engine has no arbitrary-size stage
  if (arbitrary && arb == nullptr) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  KCV_TRY(check_ptr_stride(src, src_stride, height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, height));
  KCV_TRY(check_image_size(width, height));
  if (width == 0 || height == 0 || channels == 0) return KLEIDICV_OK;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  if (arbitrary) {
    // src/filters/gaussian_blur_arbitrary_neon.cpp:395-430: the sigma
    // half-kernel (sigma 0 derived from the size), or a copy when the
    // centre tap would be 256
    uint8_t half[128];
    const size_t half_size = kernel_width / 2 + 1;
    if (!gaussian_half_kernel(half, half_size, sigma_x)) {
      *act = Action::kCopy;
      return KLEIDICV_OK;
    }
    *arb = GaussArbArgs{};
    arb->src = src;
    arb->src_stride = src_stride;
    arb->dst = dst;
    arb->dst_stride = dst_stride;
    arb->width = static_cast<int>(width);
    arb->height = static_cast<int>(height);
    arb->channels = static_cast<int>(channels);
    arb->ksize = static_cast<int>(kernel_width);
    for (size_t i = 0; i < half_size; ++i)
      arb->taps[i] = arb->taps[kernel_width - 1 - i] = half[i];
    *act = Action::kLaunchArbitrary;
    return KLEIDICV_OK;
  }

  *a = SepArgs{};
  a->src = src;
  a->src_stride = src_stride;
  a->dst = dst;
  a->dst_stride = dst_stride;
  a->width = static_cast<int>(width);
  a->height = static_cast<int>(height);
  a->channels = static_cast<int>(channels);
  a->border = static_cast<int>(border_type);
  a->ksize = static_cast<int>(kernel_width);
  if (sigma_x == 0.0f && kernel_width <= 9) {
    static const int kTaps[4][9] = {{1, 2, 1},
                                    {1, 4, 6, 4, 1},
                                    {2, 7, 14, 18, 14, 7, 2},
                                    {4, 13, 30, 51, 60, 51, 30, 13, 4}};
    static const int kShift[4] = {4, 8, 12, 16};
    const int k = static_cast<int>(kernel_width / 2) - 1;
    a->mode = kSepBinomial;
    a->shift = kShift[k];
    for (size_t i = 0; i < kernel_width; ++i) a->wx[i] = a->wy[i] = kTaps[k][i];
  } else {
    uint8_t half[16];
    const size_t half_size = kernel_width / 2 + 1;
    if (!gaussian_half_kernel(half, half_size, sigma_x)) {
      *act = Action::kCopy;  // gaussian_blur_fixed_neon.cpp:565-573
      return KLEIDICV_OK;
    }
    a->mode = kSepSigma;
    for (size_t i = 0; i < half_size; ++i)
      a->wx[i] = a->wy[i] = a->wx[kernel_width - 1 - i] = a->wy[kernel_width - 1 - i] =
          half[i];
  }
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// ---------------------------------------------------------------------------
// Blur and downsample.  src/filters/blur_and_downsample_api.cpp:18-30
// (size and channels, then the fixed border types), then
// src/filters/blur_and_downsample_neon.cpp:358-360 (pointers with the
// source and destination row counts, then the image size).
kleidicv_error_t prepare_blur_down(const uint8_t *src, size_t src_stride, size_t width,
                                   size_t height, uint8_t *dst, size_t dst_stride,
                                   size_t channels, kleidicv_border_type_t border_type,
                                   SepArgs *a, Action *act) {
  *act = Action::kNone;
  if (width < 4 || height < 4 || channels < 1 || channels > KLEIDICV_MAXIMUM_CHANNEL_COUNT)
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (!is_filter_border(border_type)) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  KCV_TRY(check_ptr_stride(src, src_stride, height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, (height + 1) / 2));
  KCV_TRY(check_image_size(width, height));
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  *a = SepArgs{};
  a->src = src;
  a->src_stride = src_stride;
  a->dst = dst;
  a->dst_stride = dst_stride;
  a->width = static_cast<int>(width);
  a->height = static_cast<int>(height);
  a->channels = static_cast<int>(channels);
  a->border = static_cast<int>(border_type);
  a->ksize = 5;
  a->mode = kSepBinomial;
  a->shift = 8;
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// Separable filter.  src/filters/separable_filter_2d_api.cpp:12-31,
// src/filters/separable_filter_2d_neon.cpp:223-236.
kleidicv_error_t prepare_separable(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                   size_t dst_stride, size_t width, size_t height,
                                   size_t channels, const uint8_t *kernel_x,
                                   size_t kernel_width, const uint8_t *kernel_y,
                                   size_t kernel_height, kleidicv_border_type_t border_type,
                                   SepArgs *a, Action *act) {
  *act = Action::kNone;
  // 5x5 as the reference; 3x3 and 7x7 extend it with the same arithmetic.
  if (kernel_width != kernel_height ||
      !(kernel_width == 3 || kernel_width == 5 || kernel_width == 7))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (width < kernel_width - 1 || height < kernel_width - 1)
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (!is_filter_border(border_type)) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  KCV_TRY(check_ptr_stride(src, src_stride, height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, height));
  KCV_TRY(check_image_size(width, height));
  if (kernel_x == nullptr || kernel_y == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  if (channels > KLEIDICV_MAXIMUM_CHANNEL_COUNT) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (width == 0 || height == 0 || channels == 0) return KLEIDICV_OK;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  *a = SepArgs{};
  a->src = src;
  a->src_stride = src_stride;
  a->dst = dst;
  a->dst_stride = dst_stride;
  a->width = static_cast<int>(width);
  a->height = static_cast<int>(height);
  a->channels = static_cast<int>(channels);
  a->border = static_cast<int>(border_type);
  a->ksize = static_cast<int>(kernel_width);
  a->mode = kSepSaturate;
  uint8_t kx[7], ky[7];
  KCV_TRY(read_small(kx, kernel_x, kernel_width));
  KCV_TRY(read_small(ky, kernel_y, kernel_width));
  for (size_t i = 0; i < kernel_width; ++i) {
    a->wx[i] = kx[i];
    a->wy[i] = ky[i];
  }
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// ---------------------------------------------------------------------------
// Median blur.  src/filters/median_blur_api.cpp:62-95,
// include/kleidicv/filters/median_blur.h:149-195.
kleidicv_error_t prepare_median(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                size_t dst_stride, size_t width, size_t height,
                                size_t channels, size_t kernel_width, size_t kernel_height,
                                kleidicv_border_type_t border_type, MedianArgs *a,
                                Action *act) {
  *act = Action::kNone;
  KCV_TRY(check_ptr_stride(src, src_stride, height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, height));
  KCV_TRY(check_image_size(width, height));
  const bool size_ok = kernel_width == kernel_height && kernel_width >= 3 &&
                       kernel_width <= 255 && (kernel_width % 2) != 0;
  if (!(src != dst && channels <= KLEIDICV_MAXIMUM_CHANNEL_COUNT &&
        height >= kernel_height - 1 && width >= kernel_width - 1 && size_ok &&
        is_filter_border(border_type)))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (width == 0 || height == 0 || channels == 0) return KLEIDICV_OK;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  *a = MedianArgs{};
  a->src = src;
  a->src_stride = src_stride;
  a->dst = dst;
  a->dst_stride = dst_stride;
  a->width = static_cast<int>(width);
  a->height = static_cast<int>(height);
  a->channels = static_cast<int>(channels);
  a->border = static_cast<int>(border_type);
  a->ksize = static_cast<int>(kernel_width);
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// ---------------------------------------------------------------------------
// Resize.  src/resize/resize_linear_api.cpp:82-247,
// include/kleidicv/resize/resize_linear.h:16-65.
kleidicv_error_t prepare_resize(const uint8_t *src, size_t src_stride, size_t src_width,
                                size_t src_height, uint8_t *dst, size_t dst_stride,
                                size_t dst_width, size_t dst_height, size_t channels,
                                ResizeArgs *a, Action *act) {
  *act = Action::kNone;
  if (!resize_is_implemented(src_width, src_height, dst_width, dst_height, channels))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  KCV_TRY(check_ptr_stride(src, src_stride, src_height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, dst_height));
  if (src_width == 0 || src_height == 0) return KLEIDICV_OK;
  if (dst_width == 0 || dst_height == 0) return KLEIDICV_OK;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  *a = ResizeArgs{};
  a->src = src;
  a->dst = dst;
  a->src_width = static_cast<int>(src_width);
  a->src_height = static_cast<int>(src_height);
  a->dst_width = static_cast<int>(dst_width);
  a->dst_height = static_cast<int>(dst_height);
  a->channels = static_cast<int>(channels);
  a->src_stride = src_stride;
  a->dst_stride = dst_stride;
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// ---------------------------------------------------------------------------
// Remap s16point5.  src/transform/remap_s16point5_neon.cpp:1212-1235,
// include/kleidicv/transform/remap.h:35-53.
kleidicv_error_t prepare_remap_s16point5(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height, size_t channels,
    const int16_t *mapxy, size_t mapxy_stride, const uint16_t *mapfrac,
    size_t mapfrac_stride, kleidicv_border_type_t border_type, const uint8_t *border_value,
    RemapS16Args *a, Action *act) {
  *act = Action::kNone;
  KCV_TRY(check_ptr_stride(src, src_stride, src_height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, dst_height));
  KCV_TRY(check_ptr_stride(mapxy, mapxy_stride, dst_height));
  KCV_TRY(check_ptr_stride(mapfrac, mapfrac_stride, dst_height));
  KCV_TRY(check_image_size(src_width, src_height));
  KCV_TRY(check_image_size(dst_width, dst_height));
  if (border_type == KLEIDICV_BORDER_TYPE_CONSTANT && border_value == nullptr)
    return KLEIDICV_ERROR_NULL_POINTER;
  if (!((channels == 1 || channels == 4) && src_stride <= 65535 / channels &&
        dst_width >= 8 && src_width <= 32768 && src_height <= 32768 &&
        (border_type == KLEIDICV_BORDER_TYPE_REPLICATE ||
         border_type == KLEIDICV_BORDER_TYPE_CONSTANT)))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (dst_height == 0) return KLEIDICV_OK;
  // An empty source leaves no pixel to clamp to (the reference reads out of
  // bounds there); declined as remap_f32_is_implemented (remap.h:55-74)
  // declines it.
  if (src_width == 0 || src_height == 0) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  *a = RemapS16Args{};
  a->src = src;
  a->dst = dst;
  a->mapxy = mapxy;
  a->mapfrac = mapfrac;
  a->src_width = static_cast<int>(src_width);
  a->src_height = static_cast<int>(src_height);
  a->dst_width = static_cast<int>(dst_width);
  a->dst_height = static_cast<int>(dst_height);
  a->channels = static_cast<int>(channels);
  a->src_stride = src_stride;
  a->dst_stride = dst_stride;
  a->mapxy_stride = mapxy_stride;
  a->mapfrac_stride = mapfrac_stride;
  a->border = static_cast<int>(border_type);
  if (border_type == KLEIDICV_BORDER_TYPE_CONSTANT)
    KCV_TRY(read_small(a->border_value, border_value, channels));
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// ---------------------------------------------------------------------------
// Remap f32.  src/transform/remap_f32_neon.cpp:183-209,
// include/kleidicv/transform/remap.h:55-74 (channels extended to 4).
kleidicv_error_t prepare_remap_f32(const uint8_t *src, size_t src_stride, size_t src_width,
                                   size_t src_height, uint8_t *dst, size_t dst_stride,
                                   size_t dst_width, size_t dst_height, size_t channels,
                                   const float *mapx, size_t mapx_stride, const float *mapy,
                                   size_t mapy_stride,
                                   kleidicv_interpolation_type_t interpolation,
                                   kleidicv_border_type_t border_type,
                                   const uint8_t *border_value, RemapF32Args *a,
                                   Action *act) {
  *act = Action::kNone;
  KCV_TRY(check_ptr_stride(src, src_stride, src_height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, dst_height));
  KCV_TRY(check_ptr_stride(mapx, mapx_stride, dst_height));
  KCV_TRY(check_ptr_stride(mapy, mapy_stride, dst_height));
  KCV_TRY(check_image_size(src_width, src_height));
  KCV_TRY(check_image_size(dst_width, dst_height));
  if (border_type == KLEIDICV_BORDER_TYPE_CONSTANT && border_value == nullptr)
    return KLEIDICV_ERROR_NULL_POINTER;
  if (!(src_stride <= 0xFFFFFFFFull && dst_width >= 4 && src_width < (1ull << 24) &&
        src_height < (1ull << 24) && dst_width < (1ull << 24) &&
        dst_height < (1ull << 24) && src_width > 0 && src_height > 0 &&
        (border_type == KLEIDICV_BORDER_TYPE_REPLICATE ||
         border_type == KLEIDICV_BORDER_TYPE_CONSTANT) &&
        channels >= 1 && channels <= 4 &&
        (interpolation == KLEIDICV_INTERPOLATION_LINEAR ||
         interpolation == KLEIDICV_INTERPOLATION_NEAREST)))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (dst_height == 0) return KLEIDICV_OK;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  *a = RemapF32Args{};
  a->src = src;
  a->dst = dst;
  a->mapx = mapx;
  a->mapy = mapy;
  a->src_width = static_cast<int>(src_width);
  a->src_height = static_cast<int>(src_height);
  a->dst_width = static_cast<int>(dst_width);
  a->dst_height = static_cast<int>(dst_height);
  a->channels = static_cast<int>(channels);
  a->src_stride = src_stride;
  a->dst_stride = dst_stride;
  a->mapx_stride = mapx_stride;
  a->mapy_stride = mapy_stride;
  a->interpolation = static_cast<int>(interpolation);
  a->border = static_cast<int>(border_type);
  if (border_type == KLEIDICV_BORDER_TYPE_CONSTANT)
    KCV_TRY(read_small(a->border_value, border_value, channels));
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// ---------------------------------------------------------------------------
// Warp perspective.  src/transform/warp_perspective_api.cpp:17-34,
// src/transform/warp_perspective_neon.cpp:158-193, transform_common.h:17-43.
kleidicv_error_t prepare_warp(const uint8_t *src, size_t src_stride, size_t src_width,
                              size_t src_height, uint8_t *dst, size_t dst_stride,
                              size_t dst_width, size_t dst_height,
                              const float transformation[9], size_t channels,
                              kleidicv_interpolation_type_t interpolation,
                              kleidicv_border_type_t border_type,
                              const uint8_t *border_value, WarpArgs *a, Action *act) {
  *act = Action::kNone;
  if (!(dst_width >= 8 &&
        (interpolation == KLEIDICV_INTERPOLATION_NEAREST ||
         interpolation == KLEIDICV_INTERPOLATION_LINEAR) &&
        (border_type == KLEIDICV_BORDER_TYPE_REPLICATE ||
         border_type == KLEIDICV_BORDER_TYPE_CONSTANT) &&
        channels >= 1 && channels <= 4))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  KCV_TRY(check_ptr_stride(src, src_stride, src_height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, dst_height));
  if (transformation == nullptr) return KLEIDICV_ERROR_NULL_POINTER;
  KCV_TRY(check_image_size(src_width, src_height));
  KCV_TRY(check_image_size(dst_width, dst_height));
  if (border_type == KLEIDICV_BORDER_TYPE_CONSTANT && border_value == nullptr)
    return KLEIDICV_ERROR_NULL_POINTER;
  if (src_width >= (1ull << 24) || src_height >= (1ull << 24) ||
      dst_width >= (1ull << 24) || dst_height >= (1ull << 24) ||
      src_stride >= (1ull << 32) || src_width == 0 || src_height == 0)
    return KLEIDICV_ERROR_RANGE;
  if (dst_height == 0) return KLEIDICV_OK;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  *a = WarpArgs{};
  a->src = src;
  a->dst = dst;
  a->src_width = static_cast<int>(src_width);
  a->src_height = static_cast<int>(src_height);
  a->dst_width = static_cast<int>(dst_width);
  a->dst_height = static_cast<int>(dst_height);
  a->channels = static_cast<int>(channels);
  a->src_stride = src_stride;
  a->dst_stride = dst_stride;
  a->interpolation = static_cast<int>(interpolation);
  a->border = static_cast<int>(border_type);
  KCV_TRY(read_small(a->t, transformation, sizeof(a->t)));
  if (border_type == KLEIDICV_BORDER_TYPE_CONSTANT)
    KCV_TRY(read_small(a->border_value, border_value, channels));
  // weight_may_be_zero (transform_common.h:17-43): T6*x + T7*y + T8 parses
  // as (T6*x + T7*y) + T8 and clang fuses the first product into the first
  // sum.
  auto weight = [&](float x, float y) { return std::fmaf(a->t[6], x, a->t[7] * y) + a->t[8]; };
  const float x1 = static_cast<float>(dst_width - 1);
  const float y1 = static_cast<float>(dst_height - 1);
  float mn = weight(0.0f, 0.0f), mx = mn;
  for (float w : {weight(x1, 0.0f), weight(0.0f, y1), weight(x1, y1)}) {
    mn = w < mn ? w : mn;
    mx = w > mx ? w : mx;
  }
  a->check_zero = (mn * mx <= 0.0f) ? 1 : 0;
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// ---------------------------------------------------------------------------
// NV12 / NV21.  src/conversions/yuv420sp_to_rgb_neon.cpp:75-160.
kleidicv_error_t prepare_yuv_sp(const uint8_t *src_y, size_t src_y_stride,
                                const uint8_t *src_uv, size_t src_uv_stride, uint8_t *dst,
                                size_t dst_stride, size_t width, size_t height,
                                kleidicv_color_conversion_t color_format, YuvArgs *a,
                                Action *act) {
  *act = Action::kNone;
  switch (color_format) {
    case KLEIDICV_NV12_TO_RGB: case KLEIDICV_NV12_TO_BGR:
    case KLEIDICV_NV12_TO_RGBA: case KLEIDICV_NV12_TO_BGRA:
    case KLEIDICV_NV21_TO_RGB: case KLEIDICV_NV21_TO_BGR:
    case KLEIDICV_NV21_TO_RGBA: case KLEIDICV_NV21_TO_BGRA:
      break;
    default:
      return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  }
  KCV_TRY(check_ptr_stride(src_y, src_y_stride, height));
  KCV_TRY(check_ptr_stride(src_uv, src_uv_stride, (height + 1) / 2));
  KCV_TRY(check_ptr_stride(dst, dst_stride, height));
  KCV_TRY(check_image_size(width, height));
  if (width == 0 || height == 0) return KLEIDICV_OK;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  *a = YuvArgs{};
  a->y = src_y;
  a->uv = src_uv;
  a->dst = dst;
  a->width = static_cast<int>(width);
  a->height = static_cast<int>(height);
  a->vu = (color_format & KLEIDICV_COLOR_CONVERSION_FLAG_VU) ? 1 : 0;
  a->bgr = (color_format & KLEIDICV_COLOR_CONVERSION_FLAG_BGR) ? 1 : 0;
  a->alpha = (color_format & KLEIDICV_COLOR_CONVERSION_FLAG_ALPHA) ? 1 : 0;
  a->y_stride = src_y_stride;
  a->uv_stride = src_uv_stride;
  a->dst_stride = dst_stride;
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// One-buffer YUV to RGB.  src/conversions/yuv_to_rgb_api.cpp:22-66 routes on
// the base format: 4:4:4 (format first, then pointers and size,
// yuv444_to_rgb_sc.h:167-185), 4:2:0 planar (format first, then the source
// with (3h+1)/2 rows, yuv420p_to_rgb_sc.h:46-50,226-290), anything else to
// 4:2:2 (pointers and size, then an even width >= 2, then the format,
// yuv422_to_rgb_sc.h:398-470).
kleidicv_error_t prepare_yuv(const uint8_t *src, size_t src_stride, uint8_t *dst,
                             size_t dst_stride, size_t width, size_t height,
                             kleidicv_color_conversion_t color_format, YuvXArgs *a,
                             Action *act) {
  *act = Action::kNone;
  const unsigned fmt = static_cast<unsigned>(color_format);
  const unsigned base = fmt & KLEIDICV_COLOR_CONVERSION_YUV_FMT_MASK;
  const unsigned flags = fmt & ~KLEIDICV_COLOR_CONVERSION_YUV_FMT_MASK;
  const unsigned bgr_alpha = KLEIDICV_COLOR_CONVERSION_FLAG_BGR | KLEIDICV_COLOR_CONVERSION_FLAG_ALPHA;
  int layout;
  size_t src_rows = height;
  if (base == KLEIDICV_COLOR_CONVERSION_FMT_YUV444) {
    if (flags & ~bgr_alpha) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
    layout = kYuv444;
  } else if (base == KLEIDICV_COLOR_CONVERSION_FMT_YUV420P) {
    if (flags & ~(bgr_alpha | KLEIDICV_COLOR_CONVERSION_FLAG_VU))
      return KLEIDICV_ERROR_NOT_IMPLEMENTED;
    layout = kYuv420P;
    src_rows = (height * 3 + 1) / 2;
  } else {
    layout = kYuv422;
  }
  KCV_TRY(check_ptr_stride(src, src_stride, src_rows));
  KCV_TRY(check_ptr_stride(dst, dst_stride, height));
  KCV_TRY(check_image_size(width, height));
  if (layout == kYuv422) {
    if (width < 2 || (width % 2) != 0) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
    const bool vu = flags & KLEIDICV_COLOR_CONVERSION_FLAG_VU;
    const bool cf = flags & KLEIDICV_COLOR_CONVERSION_FLAG_CHROMA_FIRST;
    if (base != KLEIDICV_COLOR_CONVERSION_FMT_YUV422 || (vu && cf) ||
        (flags & ~(bgr_alpha | KLEIDICV_COLOR_CONVERSION_FLAG_VU |
                   KLEIDICV_COLOR_CONVERSION_FLAG_CHROMA_FIRST)))
      return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  }
  if (width == 0 || height == 0) return KLEIDICV_OK;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  *a = YuvXArgs{};
  a->src = src;
  a->src_stride = src_stride;
  a->dst = dst;
  a->dst_stride = dst_stride;
  a->width = static_cast<int>(width);
  a->height = static_cast<int>(height);
  a->layout = layout;
  a->bgr = (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_BGR) ? 1 : 0;
  a->alpha = (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_ALPHA) ? 1 : 0;
  const size_t row_bytes = width * (layout == kYuv420P ? 1 : layout == kYuv422 ? 2 : 3);
  a->src_extent = (src_rows - 1) * src_stride + row_bytes;
  if (layout == kYuv420P) {
    // yuv420p_to_rgb_sc.h:52-75: the first chroma plane follows the luma
    // plane; the second starts height/4 source rows later, plus half a row
    // when height % 4 is 2 or 3, and then steps by the other half first
    // when height % 4 == 2.  YV12 stores V first.
    const uint8_t *p1 = src + src_stride * height;
    const uint8_t *p2 = p1 + src_stride * (height / 4) + (width / 2) * ((height % 4) / 2);
    const int ph2 = height % 4 == 2 ? 1 : 0;
    const bool yv12 = fmt & KLEIDICV_COLOR_CONVERSION_FLAG_VU;
    a->u = yv12 ? p2 : p1;
    a->v = yv12 ? p1 : p2;
    a->u_phase = yv12 ? ph2 : 0;
    a->v_phase = yv12 ? 0 : ph2;
  } else if (layout == kYuv422) {
    // pixel pair bytes: YUYV = Y0 U Y1 V, UYVY = U Y0 V Y1, YVYU = Y0 V Y1 U
    if (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_CHROMA_FIRST) {
      a->y_off = 1; a->u_off = 0; a->v_off = 2;
    } else if (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_VU) {
      a->y_off = 0; a->u_off = 3; a->v_off = 1;
    } else {
      a->y_off = 0; a->u_off = 1; a->v_off = 3;
    }
  }
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// RGB to YUV, one buffer.  src/conversions/rgb_to_yuv_api.cpp:20-45 routes
// 4:4:4 (pointers and size, then the format: rgb_to_yuv444_sc.h:273-321),
// 4:2:2 (pointers and size, an even width >= 2, then the format:
// rgb_to_yuv422_sc.h:240-304), and everything else to planar 4:2:0
// (pointers with (3h+1)/2 destination rows and size, then the format:
// rgb_to_yuv420p_neon.cpp:17-74).
kleidicv_error_t prepare_rgb_to_yuv(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                    size_t dst_stride, size_t width, size_t height,
                                    kleidicv_color_conversion_t color_format, RgbYuvArgs *a,
                                    Action *act) {
  *act = Action::kNone;
  const unsigned fmt = static_cast<unsigned>(color_format);
  const unsigned base = fmt & KLEIDICV_COLOR_CONVERSION_YUV_FMT_MASK;
  const unsigned flags = fmt & ~KLEIDICV_COLOR_CONVERSION_YUV_FMT_MASK;
  const unsigned bgr_alpha = KLEIDICV_COLOR_CONVERSION_FLAG_BGR | KLEIDICV_COLOR_CONVERSION_FLAG_ALPHA;
  int layout;
  size_t dst_rows = height;
  if (base == KLEIDICV_COLOR_CONVERSION_FMT_YUV444) {
    layout = kYuv444;
  } else if (base == KLEIDICV_COLOR_CONVERSION_FMT_YUV422) {
    layout = kYuv422;
  } else {
    layout = kYuv420P;
    dst_rows = (height * 3 + 1) / 2;
  }
  KCV_TRY(check_ptr_stride(src, src_stride, height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, dst_rows));
  KCV_TRY(check_image_size(width, height));
  if (layout == kYuv444) {
    if (flags & ~bgr_alpha) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  } else if (layout == kYuv422) {
    if (width < 2 || (width % 2) != 0) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
    const bool vu = flags & KLEIDICV_COLOR_CONVERSION_FLAG_VU;
    const bool cf = flags & KLEIDICV_COLOR_CONVERSION_FLAG_CHROMA_FIRST;
    if (vu && cf) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  } else {
    if (base != KLEIDICV_COLOR_CONVERSION_FMT_YUV420P ||
        (flags & ~(bgr_alpha | KLEIDICV_COLOR_CONVERSION_FLAG_VU)))
      return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  }
  if (width == 0 || height == 0) return KLEIDICV_OK;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  *a = RgbYuvArgs{};
  a->src = src;
  a->src_stride = src_stride;
  a->dst = dst;
  a->dst_stride = dst_stride;
  a->width = static_cast<int>(width);
  a->height = static_cast<int>(height);
  a->layout = layout;
  a->scn = (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_ALPHA) ? 4 : 3;
  a->bgr = (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_BGR) ? 1 : 0;
  a->vu = (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_VU) ? 1 : 0;
  if (layout == kYuv420P) {
    a->uv = dst + dst_stride * height;  // the chroma area follows the luma plane
    a->uv_stride = dst_stride;
  } else if (layout == kYuv422) {
    if (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_CHROMA_FIRST) {
      a->y_off = 1; a->u_off = 0; a->v_off = 2;
    } else if (a->vu) {
      a->y_off = 0; a->u_off = 3; a->v_off = 1;
    } else {
      a->y_off = 0; a->u_off = 1; a->v_off = 3;
    }
    a->vu = 0;
  }
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

// RGB to NV12 / NV21.  src/conversions/rgb_to_yuv420sp_neon.cpp:12-75:
// pointers (luma rows h, chroma rows (h+1)/2), size, then the format.
kleidicv_error_t prepare_rgb_to_yuv_sp(const uint8_t *src, size_t src_stride, uint8_t *y_dst,
                                       size_t y_stride, uint8_t *uv_dst, size_t uv_stride,
                                       size_t width, size_t height,
                                       kleidicv_color_conversion_t color_format,
                                       RgbYuvArgs *a, Action *act) {
  *act = Action::kNone;
  KCV_TRY(check_ptr_stride(src, src_stride, height));
  KCV_TRY(check_ptr_stride(y_dst, y_stride, height));
  KCV_TRY(check_ptr_stride(uv_dst, uv_stride, (height + 1) / 2));
  KCV_TRY(check_image_size(width, height));
  const unsigned fmt = static_cast<unsigned>(color_format);
  if ((fmt & KLEIDICV_COLOR_CONVERSION_YUV_FMT_MASK) != KLEIDICV_COLOR_CONVERSION_FMT_YUV420SP ||
      (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_CHROMA_FIRST))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (width == 0 || height == 0) return KLEIDICV_OK;
  if (!runtime().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  *a = RgbYuvArgs{};
  a->src = src;
  a->src_stride = src_stride;
  a->dst = y_dst;
  a->dst_stride = y_stride;
  a->uv = uv_dst;
  a->uv_stride = uv_stride;
  a->width = static_cast<int>(width);
  a->height = static_cast<int>(height);
  a->layout = kRgb420SP;
  a->scn = (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_ALPHA) ? 4 : 3;
  a->bgr = (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_BGR) ? 1 : 0;
  a->vu = (fmt & KLEIDICV_COLOR_CONVERSION_FLAG_VU) ? 1 : 0;
  *act = Action::kLaunch;
  return KLEIDICV_OK;
}

}  // namespace kcv

// Test hook (not part of the KleidiCV API): the device allocation that
// follows `after` successful ones fails with out-of-memory, once; a negative
// value disarms it.
extern "C" void kcv_test_fail_device_allocation(long after) {
  kcv::g_alloc_fail_after.store(after);
}
