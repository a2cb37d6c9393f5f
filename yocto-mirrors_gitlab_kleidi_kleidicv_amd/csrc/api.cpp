// C ABI of kleidicv_hip: argument validation in the reference's order and
// with its error codes, host/device pointer handling, and dispatch to the
// HIP launchers.
//
// Validation order per entry point follows the reference's API files
// (cited at each function); the first failing check decides the code.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "kcv_internal.h"
#include "kleidicv_hip/kleidicv.h"

namespace {

using kcv::Border;

// ---------------------------------------------------------------------------
// Runtime context: one non-blocking stream per process on the current
// device, created at first use.
struct Context {
  hipStream_t stream = nullptr;
  bool ok = false;
};

Context &context() {
  static Context ctx;
  static std::once_flag once;
  std::call_once(once, [] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return;
    if (hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking) != hipSuccess)
      return;
    ctx.ok = true;
  });
  return ctx;
}

kleidicv_error_t from_hip(hipError_t e) {
  if (e == hipSuccess) return KLEIDICV_OK;
  if (std::getenv("KLEIDICV_HIP_VERBOSE"))
    std::fprintf(stderr, "kleidicv_hip: %s\n", hipGetErrorString(e));
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return KLEIDICV_ERROR_ALLOCATION;
  return KLEIDICV_ERROR_NOT_IMPLEMENTED;
}

// The device API launches on the caller's stream as given; NULL is HIP's
// null stream, as everywhere in HIP.
hipStream_t pick_stream(void *stream) { return static_cast<hipStream_t>(stream); }

// ---------------------------------------------------------------------------
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

#define KCV_TRY(expr)                          \
  do {                                         \
    kleidicv_error_t kcv_e = (expr);           \
    if (kcv_e != KLEIDICV_OK) return kcv_e;    \
  } while (0)

bool is_filter_border(kleidicv_border_type_t b) {
  return b == KLEIDICV_BORDER_TYPE_REPLICATE || b == KLEIDICV_BORDER_TYPE_REFLECT ||
         b == KLEIDICV_BORDER_TYPE_WRAP || b == KLEIDICV_BORDER_TYPE_REVERSE;
}

// ---------------------------------------------------------------------------
// Host/device memory handling for the drop-in entry points.  Pointers that
// HIP knows as device or managed memory are used in place; anything else is
// host memory and is staged through stream-ordered device allocations.
bool is_device_pointer(const void *p) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

size_t extent(size_t stride, size_t row_bytes, size_t rows) {
  return rows == 0 ? 0 : (rows - 1) * stride + row_bytes;
}

// One staged buffer: a host range mirrored in device memory.
struct Staged {
  const void *host = nullptr;  // null when the caller's pointer is used
  void *dev = nullptr;
  size_t bytes = 0;
};

class Session {
 public:
  explicit Session(hipStream_t s) : stream_(s) {}
  ~Session() {
    for (int i = 0; i < n_; ++i)
      if (bufs_[i].dev && bufs_[i].host) (void)hipFreeAsync(bufs_[i].dev, stream_);
  }

  // Device view of an input range; copies it in when it is host memory.
  template <typename T>
  kleidicv_error_t input(const T *p, size_t bytes, const T **dev) {
    if (bytes == 0 || is_device_pointer(p)) {
      *dev = p;
      return KLEIDICV_OK;
    }
    void *d = nullptr;
    KCV_TRY(from_hip(hipMallocAsync(&d, bytes, stream_)));
    bufs_[n_++] = Staged{p, d, bytes};
    KCV_TRY(from_hip(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, stream_)));
    *dev = static_cast<const T *>(d);
    return KLEIDICV_OK;
  }

  // Device view of an output range; copied back by finish().  When rows are
  // padded (stride > row bytes) the range is first filled from the host so
  // the padding bytes keep the caller's values.
  template <typename T>
  kleidicv_error_t output(T *p, size_t stride, size_t row_bytes, size_t rows,
                          T **dev) {
    const size_t bytes = extent(stride, row_bytes, rows);
    if (bytes == 0 || is_device_pointer(p)) {
      *dev = p;
      return KLEIDICV_OK;
    }
    void *d = nullptr;
    KCV_TRY(from_hip(hipMallocAsync(&d, bytes, stream_)));
    bufs_[n_] = Staged{p, d, bytes};
    out_[n_] = true;
    ++n_;
    if (rows > 1 && stride != row_bytes)
      KCV_TRY(from_hip(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, stream_)));
    *dev = static_cast<T *>(d);
    return KLEIDICV_OK;
  }

  kleidicv_error_t scratch(size_t bytes, void **dev) {
    *dev = nullptr;
    if (bytes == 0) return KLEIDICV_OK;
    KCV_TRY(from_hip(hipMallocAsync(dev, bytes, stream_)));
    bufs_[n_++] = Staged{*dev, *dev, bytes};
    return KLEIDICV_OK;
  }

  kleidicv_error_t finish() {
    for (int i = 0; i < n_; ++i)
      if (out_[i])
        KCV_TRY(from_hip(hipMemcpyAsync(const_cast<void *>(bufs_[i].host),
                                        bufs_[i].dev, bufs_[i].bytes,
                                        hipMemcpyDeviceToHost, stream_)));
    return from_hip(hipStreamSynchronize(stream_));
  }

 private:
  hipStream_t stream_;
  Staged bufs_[8];
  bool out_[8] = {};
  int n_ = 0;
};

// ---------------------------------------------------------------------------
// Gaussian blur.  src/filters/gaussian_blur_api.cpp:21-53,
// include/kleidicv/filters/gaussian_blur.h:38-95.
bool gaussian_is_implemented(size_t w, size_t h, size_t kw, size_t kh, float sx,
                             float sy, size_t c, kleidicv_border_type_t b) {
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
// The reference computes `hk * multiplier - error` as one contracted
// expression, hence the explicit fused multiply-subtract.
bool gaussian_half_kernel(uint8_t *half, size_t half_size, float sigma) {
  const size_t mid = half_size - 1;
  const size_t ksize = mid * 2 + 1;
  if (sigma == 0.0f) sigma = static_cast<float>(ksize) * 0.15f + 0.35f;
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

kleidicv_error_t gaussian_impl(const uint8_t *src, size_t src_stride, uint8_t *dst,
                               size_t dst_stride, size_t width, size_t height,
                               size_t channels, size_t kernel_width,
                               size_t kernel_height, float sigma_x, float sigma_y,
                               kleidicv_border_type_t border_type, void *stream,
                               bool sync) {
  if (!is_filter_border(border_type)) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (!gaussian_is_implemented(width, height, kernel_width, kernel_height, sigma_x,
                               sigma_y, channels, border_type))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  // The arbitrary-size path (gaussian_blur_arbitrary_neon.cpp) is not built.
  if (!(kernel_width <= 9 || kernel_width == 15 || kernel_width == 21))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  KCV_TRY(check_ptr_stride(src, src_stride, height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, height));
  KCV_TRY(check_image_size(width, height));
  if (width == 0 || height == 0 || channels == 0) return KLEIDICV_OK;
  if (!context().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  kcv::SepArgs a{};
  a.width = static_cast<int>(width);
  a.height = static_cast<int>(height);
  a.channels = static_cast<int>(channels);
  a.border = static_cast<int>(border_type);
  a.ksize = static_cast<int>(kernel_width);
  bool copy = false;
  if (sigma_x == 0.0f && kernel_width <= 9) {
    static const int kTaps[4][9] = {{1, 2, 1},
                                    {1, 4, 6, 4, 1},
                                    {2, 7, 14, 18, 14, 7, 2},
                                    {4, 13, 30, 51, 60, 51, 30, 13, 4}};
    static const int kShift[4] = {4, 8, 12, 16};
    const int k = static_cast<int>(kernel_width / 2) - 1;
    a.mode = kcv::kSepBinomial;
    a.shift = kShift[k];
    for (size_t i = 0; i < kernel_width; ++i) a.wx[i] = a.wy[i] = kTaps[k][i];
  } else {
    uint8_t half[16];
    const size_t half_size = kernel_width / 2 + 1;
    if (!gaussian_half_kernel(half, half_size, sigma_x)) {
      copy = true;
    } else {
      a.mode = kcv::kSepSigma;
      for (size_t i = 0; i < half_size; ++i)
        a.wx[i] = a.wy[i] = a.wx[kernel_width - 1 - i] = a.wy[kernel_width - 1 - i] =
            half[i];
    }
  }

  hipStream_t s = sync ? context().stream : pick_stream(stream);
  const size_t row_bytes = width * channels;
  if (!sync) {
    a.src = src;
    a.src_stride = src_stride;
    a.dst = dst;
    a.dst_stride = dst_stride;
    if (copy)
      return from_hip(hipMemcpy2DAsync(dst, dst_stride, src, src_stride, row_bytes,
                                       height, hipMemcpyDefault, s));
    return from_hip(kcv::launch_separable(a, s));
  }
  Session ss(s);
  const uint8_t *dsrc;
  uint8_t *ddst;
  KCV_TRY(ss.input(src, extent(src_stride, row_bytes, height), &dsrc));
  KCV_TRY(ss.output(dst, dst_stride, row_bytes, height, &ddst));
  if (copy) {
    KCV_TRY(from_hip(hipMemcpy2DAsync(ddst, dst_stride, dsrc, src_stride, row_bytes,
                                      height, hipMemcpyDefault, s)));
  } else {
    a.src = dsrc;
    a.src_stride = src_stride;
    a.dst = ddst;
    a.dst_stride = dst_stride;
    KCV_TRY(from_hip(kcv::launch_separable(a, s)));
  }
  return ss.finish();
}

// ---------------------------------------------------------------------------
// Separable filter.  src/filters/separable_filter_2d_api.cpp:12-31,
// src/filters/separable_filter_2d_neon.cpp:223-236.
kleidicv_error_t separable_impl(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                size_t dst_stride, size_t width, size_t height,
                                size_t channels, const uint8_t *kernel_x,
                                size_t kernel_width, const uint8_t *kernel_y,
                                size_t kernel_height,
                                kleidicv_border_type_t border_type, void *stream,
                                bool sync) {
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
  if (!context().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  kcv::SepArgs a{};
  a.width = static_cast<int>(width);
  a.height = static_cast<int>(height);
  a.channels = static_cast<int>(channels);
  a.border = static_cast<int>(border_type);
  a.ksize = static_cast<int>(kernel_width);
  a.mode = kcv::kSepSaturate;
  // The taps are read on the host: for the device API they must therefore
  // be host memory, as in the reference's signature.
  uint8_t kx[7], ky[7];
  hipStream_t s = sync ? context().stream : pick_stream(stream);
  if (is_device_pointer(kernel_x) || is_device_pointer(kernel_y)) {
    KCV_TRY(from_hip(hipMemcpy(kx, kernel_x, kernel_width, hipMemcpyDefault)));
    KCV_TRY(from_hip(hipMemcpy(ky, kernel_y, kernel_width, hipMemcpyDefault)));
  } else {
    std::memcpy(kx, kernel_x, kernel_width);
    std::memcpy(ky, kernel_y, kernel_width);
  }
  for (size_t i = 0; i < kernel_width; ++i) {
    a.wx[i] = kx[i];
    a.wy[i] = ky[i];
  }
  const size_t row_bytes = width * channels;
  if (!sync) {
    a.src = src;
    a.src_stride = src_stride;
    a.dst = dst;
    a.dst_stride = dst_stride;
    return from_hip(kcv::launch_separable(a, s));
  }
  Session ss(s);
  const uint8_t *dsrc;
  uint8_t *ddst;
  KCV_TRY(ss.input(src, extent(src_stride, row_bytes, height), &dsrc));
  KCV_TRY(ss.output(dst, dst_stride, row_bytes, height, &ddst));
  a.src = dsrc;
  a.src_stride = src_stride;
  a.dst = ddst;
  a.dst_stride = dst_stride;
  KCV_TRY(from_hip(kcv::launch_separable(a, s)));
  return ss.finish();
}

// ---------------------------------------------------------------------------
// Median blur.  src/filters/median_blur_api.cpp:62-95,
// include/kleidicv/filters/median_blur.h:149-195.
kleidicv_error_t median_impl(const uint8_t *src, size_t src_stride, uint8_t *dst,
                             size_t dst_stride, size_t width, size_t height,
                             size_t channels, size_t kernel_width,
                             size_t kernel_height, kleidicv_border_type_t border_type,
                             void *stream, bool sync) {
  KCV_TRY(check_ptr_stride(src, src_stride, height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, height));
  KCV_TRY(check_image_size(width, height));
  const bool size_ok = kernel_width == kernel_height && kernel_width >= 3 &&
                       kernel_width <= 255 && (kernel_width % 2) != 0;
  if (!(src != dst && channels <= KLEIDICV_MAXIMUM_CHANNEL_COUNT &&
        height >= kernel_height - 1 && width >= kernel_width - 1 && size_ok &&
        is_filter_border(border_type)))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  // Histogram medians for kernels above 7x7 are not built.
  if (kernel_width > 7) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  if (width == 0 || height == 0 || channels == 0) return KLEIDICV_OK;
  if (!context().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  kcv::MedianArgs a{};
  a.width = static_cast<int>(width);
  a.height = static_cast<int>(height);
  a.channels = static_cast<int>(channels);
  a.border = static_cast<int>(border_type);
  a.ksize = static_cast<int>(kernel_width);
  hipStream_t s = sync ? context().stream : pick_stream(stream);
  const size_t row_bytes = width * channels;
  if (!sync) {
    a.src = src;
    a.src_stride = src_stride;
    a.dst = dst;
    a.dst_stride = dst_stride;
    return from_hip(kcv::launch_median(a, s));
  }
  Session ss(s);
  const uint8_t *dsrc;
  uint8_t *ddst;
  KCV_TRY(ss.input(src, extent(src_stride, row_bytes, height), &dsrc));
  KCV_TRY(ss.output(dst, dst_stride, row_bytes, height, &ddst));
  a.src = dsrc;
  a.src_stride = src_stride;
  a.dst = ddst;
  a.dst_stride = dst_stride;
  KCV_TRY(from_hip(kcv::launch_median(a, s)));
  return ss.finish();
}

// ---------------------------------------------------------------------------
// Resize.  src/resize/resize_linear_api.cpp:82-247,
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

kleidicv_error_t resize_impl(const uint8_t *src, size_t src_stride, size_t src_width,
                             size_t src_height, uint8_t *dst, size_t dst_stride,
                             size_t dst_width, size_t dst_height, size_t channels,
                             void *stream, bool sync) {
  if (!resize_is_implemented(src_width, src_height, dst_width, dst_height, channels))
    return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  KCV_TRY(check_ptr_stride(src, src_stride, src_height));
  KCV_TRY(check_ptr_stride(dst, dst_stride, dst_height));
  if (src_width == 0 || src_height == 0) return KLEIDICV_OK;
  if (dst_width == 0 || dst_height == 0) return KLEIDICV_OK;
  if (!context().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  kcv::ResizeArgs a{};
  a.src_width = static_cast<int>(src_width);
  a.src_height = static_cast<int>(src_height);
  a.dst_width = static_cast<int>(dst_width);
  a.dst_height = static_cast<int>(dst_height);
  a.channels = static_cast<int>(channels);
  a.src_stride = src_stride;
  a.dst_stride = dst_stride;
  hipStream_t s = sync ? context().stream : pick_stream(stream);
  const size_t ws_bytes = kcv::resize_workspace_bytes(a);
  Session ss(s);
  void *ws = nullptr;
  KCV_TRY(ss.scratch(ws_bytes, &ws));
  if (!sync) {
    a.src = src;
    a.dst = dst;
    KCV_TRY(from_hip(kcv::launch_resize(a, ws, s)));
    return KLEIDICV_OK;  // Session frees the workspace in stream order
  }
  const uint8_t *dsrc;
  uint8_t *ddst;
  KCV_TRY(ss.input(src, extent(src_stride, src_width * channels, src_height), &dsrc));
  KCV_TRY(ss.output(dst, dst_stride, dst_width * channels, dst_height, &ddst));
  a.src = dsrc;
  a.dst = ddst;
  KCV_TRY(from_hip(kcv::launch_resize(a, ws, s)));
  return ss.finish();
}

// ---------------------------------------------------------------------------
// Remap s16point5.  src/transform/remap_s16point5_neon.cpp:1212-1235,
// include/kleidicv/transform/remap.h:35-53.
kleidicv_error_t remap_s16point5_impl(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value, void *stream,
    bool sync) {
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
  if (!context().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  kcv::RemapS16Args a{};
  a.src_width = static_cast<int>(src_width);
  a.src_height = static_cast<int>(src_height);
  a.dst_width = static_cast<int>(dst_width);
  a.dst_height = static_cast<int>(dst_height);
  a.channels = static_cast<int>(channels);
  a.src_stride = src_stride;
  a.dst_stride = dst_stride;
  a.mapxy_stride = mapxy_stride;
  a.mapfrac_stride = mapfrac_stride;
  a.border = static_cast<int>(border_type);
  if (border_type == KLEIDICV_BORDER_TYPE_CONSTANT) {
    if (is_device_pointer(border_value))
      KCV_TRY(from_hip(hipMemcpy(a.border_value, border_value, channels, hipMemcpyDefault)));
    else
      std::memcpy(a.border_value, border_value, channels);
  }
  hipStream_t s = sync ? context().stream : pick_stream(stream);
  if (!sync) {
    a.src = src;
    a.dst = dst;
    a.mapxy = mapxy;
    a.mapfrac = mapfrac;
    return from_hip(kcv::launch_remap_s16point5(a, s));
  }
  Session ss(s);
  KCV_TRY(ss.input(src, extent(src_stride, src_width * channels, src_height), &a.src));
  KCV_TRY(ss.input(mapxy, extent(mapxy_stride, dst_width * 4, dst_height), &a.mapxy));
  KCV_TRY(ss.input(mapfrac, extent(mapfrac_stride, dst_width * 2, dst_height), &a.mapfrac));
  KCV_TRY(ss.output(dst, dst_stride, dst_width * channels, dst_height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_remap_s16point5(a, s)));
  return ss.finish();
}

// ---------------------------------------------------------------------------
// Remap f32.  src/transform/remap_f32_neon.cpp:183-209,
// include/kleidicv/transform/remap.h:55-74 (channels extended to 4).
kleidicv_error_t remap_f32_impl(const uint8_t *src, size_t src_stride,
                                size_t src_width, size_t src_height, uint8_t *dst,
                                size_t dst_stride, size_t dst_width, size_t dst_height,
                                size_t channels, const float *mapx, size_t mapx_stride,
                                const float *mapy, size_t mapy_stride,
                                kleidicv_interpolation_type_t interpolation,
                                kleidicv_border_type_t border_type,
                                const uint8_t *border_value, void *stream, bool sync) {
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
  if (!context().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  kcv::RemapF32Args a{};
  a.src_width = static_cast<int>(src_width);
  a.src_height = static_cast<int>(src_height);
  a.dst_width = static_cast<int>(dst_width);
  a.dst_height = static_cast<int>(dst_height);
  a.channels = static_cast<int>(channels);
  a.src_stride = src_stride;
  a.dst_stride = dst_stride;
  a.mapx_stride = mapx_stride;
  a.mapy_stride = mapy_stride;
  a.interpolation = static_cast<int>(interpolation);
  a.border = static_cast<int>(border_type);
  if (border_type == KLEIDICV_BORDER_TYPE_CONSTANT) {
    if (is_device_pointer(border_value))
      KCV_TRY(from_hip(hipMemcpy(a.border_value, border_value, channels, hipMemcpyDefault)));
    else
      std::memcpy(a.border_value, border_value, channels);
  }
  hipStream_t s = sync ? context().stream : pick_stream(stream);
  if (!sync) {
    a.src = src;
    a.dst = dst;
    a.mapx = mapx;
    a.mapy = mapy;
    return from_hip(kcv::launch_remap_f32(a, s));
  }
  Session ss(s);
  KCV_TRY(ss.input(src, extent(src_stride, src_width * channels, src_height), &a.src));
  KCV_TRY(ss.input(mapx, extent(mapx_stride, dst_width * 4, dst_height), &a.mapx));
  KCV_TRY(ss.input(mapy, extent(mapy_stride, dst_width * 4, dst_height), &a.mapy));
  KCV_TRY(ss.output(dst, dst_stride, dst_width * channels, dst_height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_remap_f32(a, s)));
  return ss.finish();
}

// ---------------------------------------------------------------------------
// Warp perspective.  src/transform/warp_perspective_api.cpp:17-34,
// src/transform/warp_perspective_neon.cpp:158-193, transform_common.h:17-43.
kleidicv_error_t warp_impl(const uint8_t *src, size_t src_stride, size_t src_width,
                           size_t src_height, uint8_t *dst, size_t dst_stride,
                           size_t dst_width, size_t dst_height,
                           const float transformation[9], size_t channels,
                           kleidicv_interpolation_type_t interpolation,
                           kleidicv_border_type_t border_type,
                           const uint8_t *border_value, void *stream, bool sync) {
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
  if (!context().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  kcv::WarpArgs a{};
  a.src_width = static_cast<int>(src_width);
  a.src_height = static_cast<int>(src_height);
  a.dst_width = static_cast<int>(dst_width);
  a.dst_height = static_cast<int>(dst_height);
  a.channels = static_cast<int>(channels);
  a.src_stride = src_stride;
  a.dst_stride = dst_stride;
  a.interpolation = static_cast<int>(interpolation);
  a.border = static_cast<int>(border_type);
  if (is_device_pointer(transformation))
    KCV_TRY(from_hip(hipMemcpy(a.t, transformation, sizeof(a.t), hipMemcpyDefault)));
  else
    std::memcpy(a.t, transformation, sizeof(a.t));
  if (border_type == KLEIDICV_BORDER_TYPE_CONSTANT) {
    if (is_device_pointer(border_value))
      KCV_TRY(from_hip(hipMemcpy(a.border_value, border_value, channels, hipMemcpyDefault)));
    else
      std::memcpy(a.border_value, border_value, channels);
  }
  // weight_may_be_zero (transform_common.h:17-43): T6*x + T7*y + T8 parses
  // as (T6*x + T7*y) + T8 and clang fuses the first product into the first
  // sum.
  auto weight = [&](float x, float y) {
    return std::fmaf(a.t[6], x, a.t[7] * y) + a.t[8];
  };
  const float x1 = static_cast<float>(dst_width - 1);
  const float y1 = static_cast<float>(dst_height - 1);
  float mn = weight(0.0f, 0.0f), mx = mn;
  for (float w : {weight(x1, 0.0f), weight(0.0f, y1), weight(x1, y1)}) {
    mn = w < mn ? w : mn;
    mx = w > mx ? w : mx;
  }
  a.check_zero = (mn * mx <= 0.0f) ? 1 : 0;

  hipStream_t s = sync ? context().stream : pick_stream(stream);
  if (!sync) {
    a.src = src;
    a.dst = dst;
    return from_hip(kcv::launch_warp_perspective(a, s));
  }
  Session ss(s);
  KCV_TRY(ss.input(src, extent(src_stride, src_width * channels, src_height), &a.src));
  KCV_TRY(ss.output(dst, dst_stride, dst_width * channels, dst_height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_warp_perspective(a, s)));
  return ss.finish();
}

// ---------------------------------------------------------------------------
// NV12 / NV21.  src/conversions/yuv420sp_to_rgb_neon.cpp:75-160.
kleidicv_error_t yuv_sp_impl(const uint8_t *src_y, size_t src_y_stride,
                             const uint8_t *src_uv, size_t src_uv_stride, uint8_t *dst,
                             size_t dst_stride, size_t width, size_t height,
                             kleidicv_color_conversion_t color_format, void *stream,
                             bool sync) {
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
  if (!context().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;

  kcv::YuvArgs a{};
  a.width = static_cast<int>(width);
  a.height = static_cast<int>(height);
  a.vu = (color_format & KLEIDICV_COLOR_CONVERSION_FLAG_VU) ? 1 : 0;
  a.bgr = (color_format & KLEIDICV_COLOR_CONVERSION_FLAG_BGR) ? 1 : 0;
  a.alpha = (color_format & KLEIDICV_COLOR_CONVERSION_FLAG_ALPHA) ? 1 : 0;
  a.y_stride = src_y_stride;
  a.uv_stride = src_uv_stride;
  a.dst_stride = dst_stride;
  const size_t oc = a.alpha ? 4 : 3;
  hipStream_t s = sync ? context().stream : pick_stream(stream);
  if (!sync) {
    a.y = src_y;
    a.uv = src_uv;
    a.dst = dst;
    return from_hip(kcv::launch_yuv_sp_to_rgb(a, s));
  }
  Session ss(s);
  const size_t uv_bytes = (width + 1) / 2 * 2;
  KCV_TRY(ss.input(src_y, extent(src_y_stride, width, height), &a.y));
  KCV_TRY(ss.input(src_uv, extent(src_uv_stride, uv_bytes, (height + 1) / 2), &a.uv));
  KCV_TRY(ss.output(dst, dst_stride, width * oc, height, &a.dst));
  KCV_TRY(from_hip(kcv::launch_yuv_sp_to_rgb(a, s)));
  return ss.finish();
}

kleidicv_error_t check_mt(kleidicv_thread_multithreading mt) {
  return mt.parallel == nullptr ? KLEIDICV_ERROR_NULL_POINTER : KLEIDICV_OK;
}

// Function-pointer targets for the KLEIDICV_API_DECLARATION symbols.
kleidicv_error_t yuv_sp_entry(const uint8_t *src_y, size_t src_y_stride,
                              const uint8_t *src_uv, size_t src_uv_stride, uint8_t *dst,
                              size_t dst_stride, size_t width, size_t height,
                              kleidicv_color_conversion_t color_format) {
  return yuv_sp_impl(src_y, src_y_stride, src_uv, src_uv_stride, dst, dst_stride, width,
                     height, color_format, nullptr, true);
}

kleidicv_error_t remap_s16point5_entry(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value) {
  return remap_s16point5_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                              dst_width, dst_height, channels, mapxy, mapxy_stride,
                              mapfrac, mapfrac_stride, border_type, border_value,
                              nullptr, true);
}

kleidicv_error_t remap_f32_entry(const uint8_t *src, size_t src_stride,
                                 size_t src_width, size_t src_height, uint8_t *dst,
                                 size_t dst_stride, size_t dst_width, size_t dst_height,
                                 size_t channels, const float *mapx, size_t mapx_stride,
                                 const float *mapy, size_t mapy_stride,
                                 kleidicv_interpolation_type_t interpolation,
                                 kleidicv_border_type_t border_type,
                                 const uint8_t *border_value) {
  return remap_f32_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                        dst_width, dst_height, channels, mapx, mapx_stride, mapy,
                        mapy_stride, interpolation, border_type, border_value, nullptr,
                        true);
}

}  // namespace

extern "C" {

kleidicv_error_t (*kleidicv_yuv_semiplanar_to_rgb_u8)(
    const uint8_t *, size_t, const uint8_t *, size_t, uint8_t *, size_t, size_t,
    size_t, kleidicv_color_conversion_t) = &yuv_sp_entry;

kleidicv_error_t (*kleidicv_remap_s16point5_u8)(
    const uint8_t *, size_t, size_t, size_t, uint8_t *, size_t, size_t, size_t,
    size_t, const int16_t *, size_t, const uint16_t *, size_t,
    kleidicv_border_type_t, const uint8_t *) = &remap_s16point5_entry;

kleidicv_error_t (*kleidicv_remap_f32_u8)(
    const uint8_t *, size_t, size_t, size_t, uint8_t *, size_t, size_t, size_t,
    size_t, const float *, size_t, const float *, size_t,
    kleidicv_interpolation_type_t, kleidicv_border_type_t,
    const uint8_t *) = &remap_f32_entry;

kleidicv_error_t kleidicv_gaussian_blur_u8(const uint8_t *src, size_t src_stride,
                                           uint8_t *dst, size_t dst_stride,
                                           size_t width, size_t height,
                                           size_t channels, size_t kernel_width,
                                           size_t kernel_height, float sigma_x,
                                           float sigma_y,
                                           kleidicv_border_type_t border_type) {
  return gaussian_impl(src, src_stride, dst, dst_stride, width, height, channels,
                       kernel_width, kernel_height, sigma_x, sigma_y, border_type,
                       nullptr, true);
}

kleidicv_error_t kleidicv_separable_filter_2d_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type) {
  return separable_impl(src, src_stride, dst, dst_stride, width, height, channels,
                        kernel_x, kernel_width, kernel_y, kernel_height, border_type,
                        nullptr, true);
}

kleidicv_error_t kleidicv_median_blur_u8(const uint8_t *src, size_t src_stride,
                                         uint8_t *dst, size_t dst_stride, size_t width,
                                         size_t height, size_t channels,
                                         size_t kernel_width, size_t kernel_height,
                                         kleidicv_border_type_t border_type) {
  return median_impl(src, src_stride, dst, dst_stride, width, height, channels,
                     kernel_width, kernel_height, border_type, nullptr, true);
}

kleidicv_error_t kleidicv_resize_linear_u8(const uint8_t *src, size_t src_stride,
                                           size_t src_width, size_t src_height,
                                           uint8_t *dst, size_t dst_stride,
                                           size_t dst_width, size_t dst_height,
                                           size_t channels) {
  return resize_impl(src, src_stride, src_width, src_height, dst, dst_stride, dst_width,
                     dst_height, channels, nullptr, true);
}

kleidicv_error_t kleidicv_warp_perspective_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    const float transformation[9], size_t channels,
    kleidicv_interpolation_type_t interpolation, kleidicv_border_type_t border_type,
    const uint8_t *border_value) {
  return warp_impl(src, src_stride, src_width, src_height, dst, dst_stride, dst_width,
                   dst_height, transformation, channels, interpolation, border_type,
                   border_value, nullptr, true);
}

// Threaded API: the stripe scheduler is replaced by the GPU grid.
kleidicv_error_t kleidicv_thread_gaussian_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, float sigma_x, float sigma_y,
    kleidicv_border_type_t border_type, kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return kleidicv_gaussian_blur_u8(src, src_stride, dst, dst_stride, width, height,
                                   channels, kernel_width, kernel_height, sigma_x,
                                   sigma_y, border_type);
}

kleidicv_error_t kleidicv_thread_separable_filter_2d_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type, kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return kleidicv_separable_filter_2d_u8(src, src_stride, dst, dst_stride, width,
                                         height, channels, kernel_x, kernel_width,
                                         kernel_y, kernel_height, border_type);
}

kleidicv_error_t kleidicv_thread_median_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, kleidicv_border_type_t border_type,
    kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return kleidicv_median_blur_u8(src, src_stride, dst, dst_stride, width, height,
                                 channels, kernel_width, kernel_height, border_type);
}

kleidicv_error_t kleidicv_thread_resize_linear_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return kleidicv_resize_linear_u8(src, src_stride, src_width, src_height, dst,
                                   dst_stride, dst_width, dst_height, channels);
}

kleidicv_error_t kleidicv_thread_remap_s16point5_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value,
    kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return remap_s16point5_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                              dst_width, dst_height, channels, mapxy, mapxy_stride,
                              mapfrac, mapfrac_stride, border_type, border_value,
                              nullptr, true);
}

kleidicv_error_t kleidicv_thread_remap_f32_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const float *mapx, size_t mapx_stride, const float *mapy,
    size_t mapy_stride, kleidicv_interpolation_type_t interpolation,
    kleidicv_border_type_t border_type, const uint8_t *border_value,
    kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return remap_f32_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                        dst_width, dst_height, channels, mapx, mapx_stride, mapy,
                        mapy_stride, interpolation, border_type, border_value, nullptr,
                        true);
}

kleidicv_error_t kleidicv_thread_warp_perspective_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    const float transformation[9], size_t channels,
    kleidicv_interpolation_type_t interpolation, kleidicv_border_type_t border_type,
    const uint8_t *border_value, kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return kleidicv_warp_perspective_u8(src, src_stride, src_width, src_height, dst,
                                      dst_stride, dst_width, dst_height, transformation,
                                      channels, interpolation, border_type,
                                      border_value);
}

kleidicv_error_t kleidicv_thread_yuv_semiplanar_to_rgb_u8(
    const uint8_t *src_y, size_t src_y_stride, const uint8_t *src_uv,
    size_t src_uv_stride, uint8_t *dst, size_t dst_stride, size_t width,
    size_t height, kleidicv_color_conversion_t color_format,
    kleidicv_thread_multithreading mt) {
  KCV_TRY(check_mt(mt));
  return yuv_sp_impl(src_y, src_y_stride, src_uv, src_uv_stride, dst, dst_stride, width,
                     height, color_format, nullptr, true);
}

// Device-resident API.
kleidicv_error_t kleidicv_hip_gaussian_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, float sigma_x, float sigma_y,
    kleidicv_border_type_t border_type, void *stream) {
  return gaussian_impl(src, src_stride, dst, dst_stride, width, height, channels,
                       kernel_width, kernel_height, sigma_x, sigma_y, border_type,
                       stream, false);
}

kleidicv_error_t kleidicv_hip_separable_filter_2d_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type, void *stream) {
  return separable_impl(src, src_stride, dst, dst_stride, width, height, channels,
                        kernel_x, kernel_width, kernel_y, kernel_height, border_type,
                        stream, false);
}

kleidicv_error_t kleidicv_hip_median_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, kleidicv_border_type_t border_type, void *stream) {
  return median_impl(src, src_stride, dst, dst_stride, width, height, channels,
                     kernel_width, kernel_height, border_type, stream, false);
}

kleidicv_error_t kleidicv_hip_resize_linear_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, void *stream) {
  return resize_impl(src, src_stride, src_width, src_height, dst, dst_stride, dst_width,
                     dst_height, channels, stream, false);
}

kleidicv_error_t kleidicv_hip_remap_s16point5_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value, void *stream) {
  return remap_s16point5_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                              dst_width, dst_height, channels, mapxy, mapxy_stride,
                              mapfrac, mapfrac_stride, border_type, border_value, stream,
                              false);
}

kleidicv_error_t kleidicv_hip_remap_f32_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const float *mapx, size_t mapx_stride, const float *mapy,
    size_t mapy_stride, kleidicv_interpolation_type_t interpolation,
    kleidicv_border_type_t border_type, const uint8_t *border_value, void *stream) {
  return remap_f32_impl(src, src_stride, src_width, src_height, dst, dst_stride,
                        dst_width, dst_height, channels, mapx, mapx_stride, mapy,
                        mapy_stride, interpolation, border_type, border_value, stream,
                        false);
}

kleidicv_error_t kleidicv_hip_warp_perspective_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    const float transformation[9], size_t channels,
    kleidicv_interpolation_type_t interpolation, kleidicv_border_type_t border_type,
    const uint8_t *border_value, void *stream) {
  return warp_impl(src, src_stride, src_width, src_height, dst, dst_stride, dst_width,
                   dst_height, transformation, channels, interpolation, border_type,
                   border_value, stream, false);
}

kleidicv_error_t kleidicv_hip_yuv_semiplanar_to_rgb_u8(
    const uint8_t *src_y, size_t src_y_stride, const uint8_t *src_uv,
    size_t src_uv_stride, uint8_t *dst, size_t dst_stride, size_t width,
    size_t height, kleidicv_color_conversion_t color_format, void *stream) {
  return yuv_sp_impl(src_y, src_y_stride, src_uv, src_uv_stride, dst, dst_stride, width,
                     height, color_format, stream, false);
}

void *kleidicv_hip_default_stream(void) { return context().stream; }

kleidicv_error_t kleidicv_hip_synchronize(void *stream) {
  if (!context().ok) return KLEIDICV_ERROR_NOT_IMPLEMENTED;
  return from_hip(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
}

}  // extern "C"
