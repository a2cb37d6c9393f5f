// Argument validation and kernel-argument preparation shared by the
// single-image entry points (api.cpp) and the batch engine (batch_api.cpp).
//
// Each prepare_* function applies the reference's checks in the reference's
// order, returns the first failing code, and on success fills the launcher
// arguments with the caller's pointers as given.  `*act` says what remains
// to be done: nothing (empty image), a kernel launch, or a plain 2-D copy
// (Gaussian blur whose sigma is too small to blur).  Small parameter arrays
// (taps, matrices, border values) are read on the host; they may be host or
// device memory.
#ifndef KCV_PREPARE_H
#define KCV_PREPARE_H

#include <hip/hip_runtime.h>

#include "kcv_internal.h"
#include "kleidicv_hip/kleidicv.h"

namespace kcv {

enum class Action { kNone, kLaunch, kCopy, kLaunchArbitrary };

// Process-wide runtime state: the library's own non-blocking stream on the
// device current at first use, and that device's compute-unit count.
struct Runtime {
  hipStream_t stream = nullptr;
  int device = 0;
  int compute_units = 0;
  bool ok = false;
};
Runtime &runtime();

kleidicv_error_t from_hip(hipError_t e);
bool is_device_pointer(const void *p);

// Every device allocation the library makes goes through here: stream-ordered
// (hipMallocAsync on `stream`) or not (hipMalloc).  A test hook can make one
// of them fail, as the reference's tests make malloc fail
// (test/framework/wrap_malloc.cpp), to check that callers see
// KLEIDICV_ERROR_ALLOCATION and keep their state.
hipError_t device_alloc(void **p, size_t bytes, hipStream_t stream, bool stream_ordered);

#define KCV_TRY(expr)                          \
  do {                                         \
    kleidicv_error_t kcv_e = (expr);           \
    if (kcv_e != KLEIDICV_OK) return kcv_e;    \
  } while (0)

kleidicv_error_t prepare_gaussian(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                  size_t dst_stride, size_t width, size_t height,
                                  size_t channels, size_t kernel_width,
                                  size_t kernel_height, float sigma_x, float sigma_y,
                                  kleidicv_border_type_t border_type, SepArgs *a,
                                  Action *act, GaussArbArgs *arb = nullptr);

kleidicv_error_t prepare_separable(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                   size_t dst_stride, size_t width, size_t height,
                                   size_t channels, const uint8_t *kernel_x,
                                   size_t kernel_width, const uint8_t *kernel_y,
                                   size_t kernel_height, kleidicv_border_type_t border_type,
                                   SepArgs *a, Action *act);

kleidicv_error_t prepare_blur_down(const uint8_t *src, size_t src_stride, size_t width,
                                   size_t height, uint8_t *dst, size_t dst_stride,
                                   size_t channels, kleidicv_border_type_t border_type,
                                   SepArgs *a, Action *act);

kleidicv_error_t prepare_median(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                size_t dst_stride, size_t width, size_t height,
                                size_t channels, size_t kernel_width, size_t kernel_height,
                                kleidicv_border_type_t border_type, MedianArgs *a,
                                Action *act);

kleidicv_error_t prepare_resize(const uint8_t *src, size_t src_stride, size_t src_width,
                                size_t src_height, uint8_t *dst, size_t dst_stride,
                                size_t dst_width, size_t dst_height, size_t channels,
                                ResizeArgs *a, Action *act);

kleidicv_error_t prepare_remap_s16point5(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height, size_t channels,
    const int16_t *mapxy, size_t mapxy_stride, const uint16_t *mapfrac,
    size_t mapfrac_stride, kleidicv_border_type_t border_type, const uint8_t *border_value,
    RemapS16Args *a, Action *act);

kleidicv_error_t prepare_remap_f32(const uint8_t *src, size_t src_stride, size_t src_width,
                                   size_t src_height, uint8_t *dst, size_t dst_stride,
                                   size_t dst_width, size_t dst_height, size_t channels,
                                   const float *mapx, size_t mapx_stride, const float *mapy,
                                   size_t mapy_stride,
                                   kleidicv_interpolation_type_t interpolation,
                                   kleidicv_border_type_t border_type,
                                   const uint8_t *border_value, RemapF32Args *a,
                                   Action *act);

kleidicv_error_t prepare_warp(const uint8_t *src, size_t src_stride, size_t src_width,
                              size_t src_height, uint8_t *dst, size_t dst_stride,
                              size_t dst_width, size_t dst_height,
                              const float transformation[9], size_t channels,
                              kleidicv_interpolation_type_t interpolation,
                              kleidicv_border_type_t border_type,
                              const uint8_t *border_value, WarpArgs *a, Action *act);

kleidicv_error_t prepare_yuv(const uint8_t *src, size_t src_stride, uint8_t *dst,
                             size_t dst_stride, size_t width, size_t height,
                             kleidicv_color_conversion_t color_format, YuvXArgs *a,
                             Action *act);

kleidicv_error_t prepare_rgb_to_yuv(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                    size_t dst_stride, size_t width, size_t height,
                                    kleidicv_color_conversion_t color_format, RgbYuvArgs *a,
                                    Action *act);

kleidicv_error_t prepare_rgb_to_yuv_sp(const uint8_t *src, size_t src_stride, uint8_t *y_dst,
                                       size_t y_stride, uint8_t *uv_dst, size_t uv_stride,
                                       size_t width, size_t height,
                                       kleidicv_color_conversion_t color_format,
                                       RgbYuvArgs *a, Action *act);

kleidicv_error_t prepare_yuv_sp(const uint8_t *src_y, size_t src_y_stride,
                                const uint8_t *src_uv, size_t src_uv_stride, uint8_t *dst,
                                size_t dst_stride, size_t width, size_t height,
                                kleidicv_color_conversion_t color_format, YuvArgs *a,
                                Action *act);

}  // namespace kcv

#endif  // KCV_PREPARE_H
