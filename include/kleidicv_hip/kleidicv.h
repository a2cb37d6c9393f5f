/*
 * kleidicv_hip: MI355X (gfx950) implementation of KleidiCV's u8 image hot
 * path behind the reference's C ABI.
 *
 * Three layers are declared here:
 *   1. Drop-in `kleidicv_*` entry points with the reference's signatures,
 *      argument meaning and error codes.  They accept host pointers (staged
 *      through device memory) or device pointers (used in place) and return
 *      when the result is in `dst`, as the reference's CPU functions do.
 *   2. `kleidicv_thread_*` entry points with the reference's scheduler
 *      argument.  The row-stripe scheduler is replaced by the GPU grid: `mt`
 *      is validated and otherwise unused.
 *   3. `kleidicv_hip_*` device-resident entry points: device pointers, an
 *      explicit HIP stream, asynchronous.  They carry no counterpart in the
 *      reference; they exist so callers can keep images in HBM and overlap
 *      work (the batched and served paths build on them).
 *
 * Enum values are the reference's (kleidicv/include/kleidicv/ctypes.h).
 */
#ifndef KLEIDICV_HIP_KLEIDICV_H
#define KLEIDICV_HIP_KLEIDICV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ctypes.h:26-41 */
typedef enum {
  KLEIDICV_OK = 0,
  KLEIDICV_ERROR_NOT_IMPLEMENTED,
  KLEIDICV_ERROR_NULL_POINTER,
  KLEIDICV_ERROR_RANGE,
  KLEIDICV_ERROR_ALLOCATION,
  KLEIDICV_ERROR_ALIGNMENT,
  KLEIDICV_ERROR_CONTEXT_MISMATCH,
} kleidicv_error_t;

/* ctypes.h:60-79 */
typedef enum {
  KLEIDICV_BORDER_TYPE_CONSTANT,
  KLEIDICV_BORDER_TYPE_REPLICATE,
  KLEIDICV_BORDER_TYPE_REFLECT,
  KLEIDICV_BORDER_TYPE_WRAP,
  KLEIDICV_BORDER_TYPE_REVERSE,
  KLEIDICV_BORDER_TYPE_TRANSPARENT,
  KLEIDICV_BORDER_TYPE_NONE,
} kleidicv_border_type_t;

/* ctypes.h:82-87 */
typedef enum {
  KLEIDICV_INTERPOLATION_NEAREST,
  KLEIDICV_INTERPOLATION_LINEAR,
} kleidicv_interpolation_type_t;

/* ctypes.h:149-360: base format in the low nibble, modifier flags above. */
typedef enum {
  KLEIDICV_COLOR_CONVERSION_FMT_YUV444 = 0x01,
  KLEIDICV_COLOR_CONVERSION_FMT_YUV420SP = 0x02,
  KLEIDICV_COLOR_CONVERSION_FMT_YUV420P = 0x03,
  KLEIDICV_COLOR_CONVERSION_FMT_YUV422 = 0x04,
  KLEIDICV_COLOR_CONVERSION_YUV_FMT_MASK = 0x0F,
  KLEIDICV_COLOR_CONVERSION_FLAG_VU = 0x10,
  KLEIDICV_COLOR_CONVERSION_FLAG_BGR = 0x20,
  KLEIDICV_COLOR_CONVERSION_FLAG_ALPHA = 0x40,
  KLEIDICV_COLOR_CONVERSION_FLAG_CHROMA_FIRST = 0x80,
  KLEIDICV_NV12_TO_RGB = 0x02,
  KLEIDICV_NV12_TO_BGR = 0x22,
  KLEIDICV_NV12_TO_RGBA = 0x42,
  KLEIDICV_NV12_TO_BGRA = 0x62,
  KLEIDICV_NV21_TO_RGB = 0x12,
  KLEIDICV_NV21_TO_BGR = 0x32,
  KLEIDICV_NV21_TO_RGBA = 0x52,
  KLEIDICV_NV21_TO_BGRA = 0x72,
  KLEIDICV_YUV444_TO_RGB = 0x01,
  KLEIDICV_YUV444_TO_BGR = 0x21,
  KLEIDICV_YUV444_TO_RGBA = 0x41,
  KLEIDICV_YUV444_TO_BGRA = 0x61,
  KLEIDICV_IYUV_TO_RGB = 0x03,
  KLEIDICV_IYUV_TO_BGR = 0x23,
  KLEIDICV_IYUV_TO_RGBA = 0x43,
  KLEIDICV_IYUV_TO_BGRA = 0x63,
  KLEIDICV_YV12_TO_RGB = 0x13,
  KLEIDICV_YV12_TO_BGR = 0x33,
  KLEIDICV_YV12_TO_RGBA = 0x53,
  KLEIDICV_YV12_TO_BGRA = 0x73,
  KLEIDICV_YUYV_TO_RGB = 0x04,
  KLEIDICV_YUYV_TO_BGR = 0x24,
  KLEIDICV_YUYV_TO_RGBA = 0x44,
  KLEIDICV_YUYV_TO_BGRA = 0x64,
  KLEIDICV_YVYU_TO_RGB = 0x14,
  KLEIDICV_YVYU_TO_BGR = 0x34,
  KLEIDICV_YVYU_TO_RGBA = 0x54,
  KLEIDICV_YVYU_TO_BGRA = 0x74,
  KLEIDICV_UYVY_TO_RGB = 0x84,
  KLEIDICV_UYVY_TO_BGR = 0xA4,
  KLEIDICV_UYVY_TO_RGBA = 0xC4,
  KLEIDICV_UYVY_TO_BGRA = 0xE4,
  /* RGB -> YUV names: the same values, the flags describing the source */
  KLEIDICV_RGB_TO_YUV444 = 0x01,
  KLEIDICV_RGB_TO_NV12 = 0x02,
  KLEIDICV_RGB_TO_NV21 = 0x12,
  KLEIDICV_RGB_TO_IYUV = 0x03,
  KLEIDICV_RGB_TO_YV12 = 0x13,
  KLEIDICV_RGB_TO_YUYV = 0x04,
  KLEIDICV_RGB_TO_YVYU = 0x14,
  KLEIDICV_RGB_TO_UYVY = 0x84,
  KLEIDICV_BGR_TO_YUV444 = 0x21,
  KLEIDICV_BGR_TO_NV12 = 0x22,
  KLEIDICV_BGR_TO_NV21 = 0x32,
  KLEIDICV_BGR_TO_IYUV = 0x23,
  KLEIDICV_BGR_TO_YV12 = 0x33,
  KLEIDICV_BGR_TO_YUYV = 0x24,
  KLEIDICV_BGR_TO_YVYU = 0x34,
  KLEIDICV_BGR_TO_UYVY = 0xA4,
  KLEIDICV_RGBA_TO_YUV444 = 0x41,
  KLEIDICV_RGBA_TO_NV12 = 0x42,
  KLEIDICV_RGBA_TO_NV21 = 0x52,
  KLEIDICV_RGBA_TO_IYUV = 0x43,
  KLEIDICV_RGBA_TO_YV12 = 0x53,
  KLEIDICV_RGBA_TO_YUYV = 0x44,
  KLEIDICV_RGBA_TO_YVYU = 0x54,
  KLEIDICV_RGBA_TO_UYVY = 0xC4,
  KLEIDICV_BGRA_TO_YUV444 = 0x61,
  KLEIDICV_BGRA_TO_NV12 = 0x62,
  KLEIDICV_BGRA_TO_NV21 = 0x72,
  KLEIDICV_BGRA_TO_IYUV = 0x63,
  KLEIDICV_BGRA_TO_YV12 = 0x73,
  KLEIDICV_BGRA_TO_YUYV = 0x64,
  KLEIDICV_BGRA_TO_YVYU = 0x74,
  KLEIDICV_BGRA_TO_UYVY = 0xE4,
} kleidicv_color_conversion_t;

/* kleidicv.h:103-113 */
#define KLEIDICV_MAX_IMAGE_PIXELS ((1ULL << 48) - 1)
#define KLEIDICV_MAXIMUM_CHANNEL_COUNT (4)
#define KLEIDICV_RESIZE_MAX_WIDTH_OR_HEIGHT ((1ULL << 24) - 1)

/* ------------------------------------------------------------------------
 * 1. Drop-in C API
 * ------------------------------------------------------------------------ */

/* Replaces kleidicv.h:2159-2162 (src/filters/gaussian_blur_api.cpp:59-67).
 * Kernel sizes 3, 5, 7, 9 (binomial when sigma == 0), 15 and 21; borders
 * REPLICATE, REFLECT, WRAP, REVERSE. */
kleidicv_error_t kleidicv_gaussian_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, float sigma_x, float sigma_y,
    kleidicv_border_type_t border_type);

/* Replaces kleidicv.h:2091-2095 (src/filters/separable_filter_2d_api.cpp:
 * 46-54).  5x5 as the reference; 3x3 and 7x7 are extensions computed with the
 * same saturating arithmetic. */
kleidicv_error_t kleidicv_separable_filter_2d_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type);

/* Replaces kleidicv.h:1579-1581 (src/conversions/yuv_to_rgb_api.cpp:
 * 22-66): one-buffer YUV to RGB/BGR/RGBA/BGRA.  Planar 4:2:0 (I420/IYUV,
 * YV12: luma plane then two chroma planes in OpenCV's packed layout, (3h+1)/2
 * rows of src_stride), interleaved 4:2:2 (YUYV, UYVY, YVYU; even width) and
 * interleaved 4:4:4 (full-range BT.601). */
kleidicv_error_t kleidicv_yuv_to_rgb_u8(const uint8_t *src, size_t src_stride,
                                        uint8_t *dst, size_t dst_stride,
                                        size_t width, size_t height,
                                        kleidicv_color_conversion_t color_format);

/* Replaces kleidicv.h:1740-1742 (src/conversions/rgb_to_yuv_api.cpp:
 * 65-70): RGB/BGR(A) to planar 4:2:0 (I420/YV12 in OpenCV's packed layout,
 * (3h+1)/2 destination rows), interleaved 4:2:2 and 4:4:4. */
kleidicv_error_t kleidicv_rgb_to_yuv_u8(const uint8_t *src, size_t src_stride,
                                        uint8_t *dst, size_t dst_stride,
                                        size_t width, size_t height,
                                        kleidicv_color_conversion_t color_format);

/* Replaces kleidicv.h:1811-1814 (rgb_to_yuv_api.cpp:47-55): RGB/BGR(A) to
 * NV12/NV21, chroma rows of (width+1)/2 pairs. */
kleidicv_error_t kleidicv_rgb_to_yuv_semiplanar_u8(
    const uint8_t *src, size_t src_stride, uint8_t *y_dst, size_t y_stride,
    uint8_t *uv_dst, size_t uv_stride, size_t width, size_t height,
    kleidicv_color_conversion_t color_format);

/* Replaces kleidicv.h:2215-2218 (src/filters/blur_and_downsample_api.cpp:
 * 37-45): the binomial 5x5 Gaussian at even rows and columns (cv::pyrDown),
 * destination ((src_width+1)/2) x ((src_height+1)/2). */
kleidicv_error_t kleidicv_blur_and_downsample_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t channels,
    kleidicv_border_type_t border_type);

/* Replaces kleidicv.h:2271 (src/filters/median_blur_api.cpp:121-130).
 * Kernel sizes 3, 5 and 7 (the reference's sorting-network sizes). */
kleidicv_error_t kleidicv_median_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, kleidicv_border_type_t border_type);

/* Replaces kleidicv.h:2483-2487 (src/resize/resize_linear_api.cpp:269-277). */
kleidicv_error_t kleidicv_resize_linear_u8(const uint8_t *src,
                                           size_t src_stride, size_t src_width,
                                           size_t src_height, uint8_t *dst,
                                           size_t dst_stride, size_t dst_width,
                                           size_t dst_height, size_t channels);

/* Replaces kleidicv.h:2899-2904 (src/transform/warp_perspective_api.cpp:
 * 17-34).  1 channel as the reference; 2-4 channels are per-plane
 * extensions. */
kleidicv_error_t kleidicv_warp_perspective_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    const float transformation[9], size_t channels,
    kleidicv_interpolation_type_t interpolation,
    kleidicv_border_type_t border_type, const uint8_t *border_value);

/* Function-pointer data symbols, as KLEIDICV_API_DECLARATION declares them
 * (kleidicv.h:126-130); initialised to this library's implementations. */

/* kleidicv.h:1640-1645 (src/conversions/yuv_to_rgb_api.cpp:23-24) */
extern kleidicv_error_t (*kleidicv_yuv_semiplanar_to_rgb_u8)(
    const uint8_t *src_y, size_t src_y_stride, const uint8_t *src_uv,
    size_t src_uv_stride, uint8_t *dst, size_t dst_stride, size_t width,
    size_t height, kleidicv_color_conversion_t color_format);

/* kleidicv.h:2715-2721 (src/transform/remap_api.cpp:17-19) */
extern kleidicv_error_t (*kleidicv_remap_s16point5_u8)(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value);

/* kleidicv.h:2782-2789 (src/transform/remap_api.cpp:25-27).  1-2 channels
 * as the reference; 3-4 channels are per-plane extensions. */
extern kleidicv_error_t (*kleidicv_remap_f32_u8)(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const float *mapx, size_t mapx_stride, const float *mapy,
    size_t mapy_stride, kleidicv_interpolation_type_t interpolation,
    kleidicv_border_type_t border_type, const uint8_t *border_value);

/* Backend-selection twins (kleidicv.h:21-35): the reference's `_sme` names
 * choose its SME backend and keep the semantics and parameter contract of
 * the unsuffixed functions.  A GPU has one backend, so each twin runs
 * exactly the unsuffixed entry point. */

/* kleidicv.h:1583-1585 */
kleidicv_error_t kleidicv_yuv_to_rgb_u8_sme(const uint8_t *src, size_t src_stride,
                                            uint8_t *dst, size_t dst_stride,
                                            size_t width, size_t height,
                                            kleidicv_color_conversion_t color_format);
/* kleidicv.h:1744-1746 */
kleidicv_error_t kleidicv_rgb_to_yuv_u8_sme(const uint8_t *src, size_t src_stride,
                                            uint8_t *dst, size_t dst_stride,
                                            size_t width, size_t height,
                                            kleidicv_color_conversion_t color_format);
/* kleidicv.h:1816-1819 */
kleidicv_error_t kleidicv_rgb_to_yuv_semiplanar_u8_sme(
    const uint8_t *src, size_t src_stride, uint8_t *y_dst, size_t y_stride,
    uint8_t *uv_dst, size_t uv_stride, size_t width, size_t height,
    kleidicv_color_conversion_t color_format);
/* kleidicv.h:2097-2100 */
kleidicv_error_t kleidicv_separable_filter_2d_u8_sme(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type);
/* kleidicv.h:2165-2168 */
kleidicv_error_t kleidicv_gaussian_blur_u8_sme(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, float sigma_x, float sigma_y,
    kleidicv_border_type_t border_type);
/* kleidicv.h:2220-2223 */
kleidicv_error_t kleidicv_blur_and_downsample_u8_sme(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t channels,
    kleidicv_border_type_t border_type);
/* kleidicv.h:2273 */
kleidicv_error_t kleidicv_median_blur_u8_sme(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, kleidicv_border_type_t border_type);
/* kleidicv.h:2489-2492 */
kleidicv_error_t kleidicv_resize_linear_u8_sme(const uint8_t *src,
                                               size_t src_stride, size_t src_width,
                                               size_t src_height, uint8_t *dst,
                                               size_t dst_stride, size_t dst_width,
                                               size_t dst_height, size_t channels);
/* kleidicv.h:1647-1652: a function-pointer data symbol like its twin */
extern kleidicv_error_t (*kleidicv_yuv_semiplanar_to_rgb_u8_sme)(
    const uint8_t *src_y, size_t src_y_stride, const uint8_t *src_uv,
    size_t src_uv_stride, uint8_t *dst, size_t dst_stride, size_t width,
    size_t height, kleidicv_color_conversion_t color_format);

/* ------------------------------------------------------------------------
 * 2. Threaded API (kleidicv_thread/include/kleidicv_thread/kleidicv_thread.h)
 * ------------------------------------------------------------------------ */

/* kleidicv_thread.h:55-97 */
typedef kleidicv_error_t (*kleidicv_thread_callback)(unsigned begin,
                                                     unsigned end, void *data);
typedef kleidicv_error_t (*kleidicv_thread_parallel)(
    kleidicv_thread_callback callback, void *callback_data, void *parallel_data,
    unsigned task_count);
typedef struct {
  kleidicv_thread_parallel parallel;
  void *parallel_data;
} kleidicv_thread_multithreading;

/* kleidicv_thread.h:459 */
kleidicv_error_t kleidicv_thread_gaussian_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, float sigma_x, float sigma_y,
    kleidicv_border_type_t border_type, kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:467 */
kleidicv_error_t kleidicv_thread_separable_filter_2d_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type, kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:223-227 */
kleidicv_error_t kleidicv_thread_rgb_to_yuv_semiplanar_u8(
    const uint8_t *src, size_t src_stride, uint8_t *y_dst, size_t y_stride,
    uint8_t *uv_dst, size_t uv_stride, size_t width, size_t height,
    kleidicv_color_conversion_t color_format, kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:231-234 */
kleidicv_error_t kleidicv_thread_rgb_to_yuv_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, kleidicv_color_conversion_t color_format,
    kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:208-211 */
kleidicv_error_t kleidicv_thread_yuv_to_rgb_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, kleidicv_color_conversion_t color_format,
    kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:482-485 */
kleidicv_error_t kleidicv_thread_blur_and_downsample_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t channels,
    kleidicv_border_type_t border_type, kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:496 */
kleidicv_error_t kleidicv_thread_median_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, kleidicv_border_type_t border_type,
    kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:542 */
kleidicv_error_t kleidicv_thread_resize_linear_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:578 */
kleidicv_error_t kleidicv_thread_remap_s16point5_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value,
    kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:602 */
kleidicv_error_t kleidicv_thread_remap_f32_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const float *mapx, size_t mapx_stride, const float *mapy,
    size_t mapy_stride, kleidicv_interpolation_type_t interpolation,
    kleidicv_border_type_t border_type, const uint8_t *border_value,
    kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:622 */
kleidicv_error_t kleidicv_thread_warp_perspective_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    const float transformation[9], size_t channels,
    kleidicv_interpolation_type_t interpolation,
    kleidicv_border_type_t border_type, const uint8_t *border_value,
    kleidicv_thread_multithreading mt);
/* kleidicv_thread.h:215 */
kleidicv_error_t kleidicv_thread_yuv_semiplanar_to_rgb_u8(
    const uint8_t *src_y, size_t src_y_stride, const uint8_t *src_uv,
    size_t src_uv_stride, uint8_t *dst, size_t dst_stride, size_t width,
    size_t height, kleidicv_color_conversion_t color_format,
    kleidicv_thread_multithreading mt);

/* ------------------------------------------------------------------------
 * 3. Device-resident API: device pointers, `stream` is a hipStream_t (NULL =
 *    HIP's null stream), no synchronisation.  Same validation and error
 *    codes as the drop-in functions.
 * ------------------------------------------------------------------------ */

kleidicv_error_t kleidicv_hip_gaussian_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, float sigma_x, float sigma_y,
    kleidicv_border_type_t border_type, void *stream);

kleidicv_error_t kleidicv_hip_separable_filter_2d_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kernel_x,
    size_t kernel_width, const uint8_t *kernel_y, size_t kernel_height,
    kleidicv_border_type_t border_type, void *stream);

kleidicv_error_t kleidicv_hip_rgb_to_yuv_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, kleidicv_color_conversion_t color_format,
    void *stream);

kleidicv_error_t kleidicv_hip_rgb_to_yuv_semiplanar_u8(
    const uint8_t *src, size_t src_stride, uint8_t *y_dst, size_t y_stride,
    uint8_t *uv_dst, size_t uv_stride, size_t width, size_t height,
    kleidicv_color_conversion_t color_format, void *stream);

kleidicv_error_t kleidicv_hip_yuv_to_rgb_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, kleidicv_color_conversion_t color_format,
    void *stream);

kleidicv_error_t kleidicv_hip_blur_and_downsample_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t channels,
    kleidicv_border_type_t border_type, void *stream);

kleidicv_error_t kleidicv_hip_median_blur_u8(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, size_t kernel_width,
    size_t kernel_height, kleidicv_border_type_t border_type, void *stream);

kleidicv_error_t kleidicv_hip_resize_linear_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, void *stream);

kleidicv_error_t kleidicv_hip_remap_s16point5_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const int16_t *mapxy, size_t mapxy_stride,
    const uint16_t *mapfrac, size_t mapfrac_stride,
    kleidicv_border_type_t border_type, const uint8_t *border_value,
    void *stream);

kleidicv_error_t kleidicv_hip_remap_f32_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, const float *mapx, size_t mapx_stride, const float *mapy,
    size_t mapy_stride, kleidicv_interpolation_type_t interpolation,
    kleidicv_border_type_t border_type, const uint8_t *border_value,
    void *stream);

kleidicv_error_t kleidicv_hip_warp_perspective_u8(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    const float transformation[9], size_t channels,
    kleidicv_interpolation_type_t interpolation,
    kleidicv_border_type_t border_type, const uint8_t *border_value,
    void *stream);

kleidicv_error_t kleidicv_hip_yuv_semiplanar_to_rgb_u8(
    const uint8_t *src_y, size_t src_y_stride, const uint8_t *src_uv,
    size_t src_uv_stride, uint8_t *dst, size_t dst_stride, size_t width,
    size_t height, kleidicv_color_conversion_t color_format, void *stream);

/* No-allocation resize.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * validation, results and error codes as kleidicv_hip_resize_linear_u8, but
 * the coordinate tables of the generic and exact-2x/4x paths go to a fixed
 * device array of KLEIDICV_HIP_NOALLOC_SCRATCH_BYTES that each GPU holds
 * until the process ends, and no allocator is called.  The tables take
 * (dst_width * channels + dst_height) * 8 bytes; a call that needs more
 * returns KLEIDICV_ERROR_RANGE and launches nothing.  Calls on one GPU share
 * the array: calls that overlap on the device (other streams, other host
 * threads) race and must be ordered by the caller. */
#define KLEIDICV_HIP_NOALLOC_SCRATCH_BYTES ((size_t)4 << 20)
kleidicv_error_t kleidicv_hip_resize_linear_u8_noalloc(
    const uint8_t *src, size_t src_stride, size_t src_width, size_t src_height,
    uint8_t *dst, size_t dst_stride, size_t dst_width, size_t dst_height,
    size_t channels, void *stream);

/* The library's own non-blocking stream (used by the drop-in entry points),
 * and a helper that waits for a stream (NULL = HIP's null stream). */
void *kleidicv_hip_default_stream(void);
kleidicv_error_t kleidicv_hip_synchronize(void *stream);

/* ------------------------------------------------------------------------
 * 4. Batched execution.  No counterpart in the reference: it replaces many
 *    per-image calls (and the reference's per-call stripe scheduling) with
 *    one submission.  Every op keeps exactly the arithmetic, validation and
 *    error codes of its single-image entry point; only the scheduling
 *    differs.
 *
 *    A batch is a list of op descriptors with device pointers.  Ops carry a
 *    stage number: every op of stage s starts after every op of stages < s
 *    has finished (a pyramid level reads the previous level), while ops of
 *    one stage run together.  For each (stage, kernel family, kernel
 *    variant) the batch builds a device-side work list of tiles across all
 *    of its images and runs it with one persistent kernel, so a batch of
 *    hundreds of images is a handful of launches.
 * ------------------------------------------------------------------------ */

typedef enum {
  KLEIDICV_HIP_OP_GAUSSIAN_BLUR = 1,
  KLEIDICV_HIP_OP_SEPARABLE_FILTER_2D = 2,
  KLEIDICV_HIP_OP_MEDIAN_BLUR = 3,
  KLEIDICV_HIP_OP_RESIZE_LINEAR = 4,
  KLEIDICV_HIP_OP_REMAP_S16POINT5 = 5,
  KLEIDICV_HIP_OP_REMAP_F32 = 6,
  KLEIDICV_HIP_OP_WARP_PERSPECTIVE = 7,
  KLEIDICV_HIP_OP_YUV_SEMIPLANAR_TO_RGB = 8,
} kleidicv_hip_op_t;

/* One op.  Fields an op does not use are ignored.
 *   src/src_stride/src_width/src_height: source image (for NV12: the Y plane)
 *   dst/dst_stride/dst_width/dst_height: destination image; filters and
 *       NV12 use src_width/src_height for both
 *   aux0/aux0_stride: mapxy (remap s16point5), mapx (remap f32), UV plane
 *       (NV12)
 *   aux1/aux1_stride: mapfrac (remap s16point5), mapy (remap f32)
 *   kernel_x/kernel_y: separable taps (kernel_size of them, by value)
 *   transformation: warp-perspective matrix (by value)
 *   border_value: CONSTANT border value per channel (by value) */
typedef struct {
  kleidicv_hip_op_t op;
  unsigned stage;
  const uint8_t *src;
  size_t src_stride, src_width, src_height;
  uint8_t *dst;
  size_t dst_stride, dst_width, dst_height;
  size_t channels;
  size_t kernel_size;
  float sigma;
  uint8_t kernel_x[7], kernel_y[7];
  kleidicv_border_type_t border_type;
  kleidicv_interpolation_type_t interpolation;
  uint8_t border_value[4];
  const void *aux0;
  size_t aux0_stride;
  const void *aux1;
  size_t aux1_stride;
  float transformation[9];
  kleidicv_color_conversion_t color_format;
} kleidicv_hip_op_desc;

/* Scheduling parameters; results never depend on them.
 *   blocks_per_cu: persistent workgroups per compute unit for each stage
 *       kernel (1..16), capped at what that kernel keeps resident at once
 *   resize_band_rows: destination rows per resize work item (0: as many as
 *       fit the LDS window, at most 64)
 *   xcd_order: 1 = hand consecutive work items (neighbouring tiles) to the
 *       workgroups of one XCD so their halos share that XCD's L2
 *   use_graph: must be 0 (stage kernels launched directly); 1 (graph
 *       replay) returns KLEIDICV_ERROR_NOT_IMPLEMENTED */
typedef struct {
  unsigned blocks_per_cu;
  unsigned resize_band_rows;
  unsigned xcd_order;
  unsigned use_graph;
} kleidicv_hip_batch_tuning;

typedef struct kleidicv_hip_batch kleidicv_hip_batch;

/* Validates every op (as its single-image entry point would), builds the
 * work lists.  On a validation error *failed_index (if not
 * NULL) names the first failing op and no batch is created.  tuning may be
 * NULL for the defaults. */
kleidicv_error_t kleidicv_hip_batch_create(const kleidicv_hip_op_desc *ops, size_t count,
                                           const kleidicv_hip_batch_tuning *tuning,
                                           kleidicv_hip_batch **batch,
                                           size_t *failed_index);

/* Enqueues the whole batch on `stream` (a hipStream_t; NULL = null stream).
 * Asynchronous.  Runs of one batch must not overlap each other. */
kleidicv_error_t kleidicv_hip_batch_run(kleidicv_hip_batch *batch, void *stream);

/* Times each candidate tuning over `iterations` runs on `stream`, keeps the
 * fastest in the batch and reports it (and its mean time per run in
 * milliseconds) when the out-pointers are not NULL.  Synchronous. */
kleidicv_error_t kleidicv_hip_batch_tune(kleidicv_hip_batch *batch, void *stream,
                                         unsigned iterations,
                                         kleidicv_hip_batch_tuning *best, float *best_ms);

kleidicv_error_t kleidicv_hip_batch_set_tuning(kleidicv_hip_batch *batch,
                                               const kleidicv_hip_batch_tuning *tuning);
kleidicv_error_t kleidicv_hip_batch_get_tuning(const kleidicv_hip_batch *batch,
                                               kleidicv_hip_batch_tuning *tuning);

/* Work items (tiles) of the batch under its current tuning. */
size_t kleidicv_hip_batch_work_items(const kleidicv_hip_batch *batch);

/* One-shot form for callers that build a new batch per submission (the
 * serving loop): validates the ops, copies their work lists into the
 * caller's device `workspace` (through pinned staging, on `stream`) and
 * enqueues the stage kernels on `stream`.  No device allocation and no wait
 * for the device.  The workspace must hold the bytes that
 * kleidicv_hip_batch_workspace_size() reports (also returned in
 * *required_bytes; a smaller workspace gives KLEIDICV_ERROR_RANGE and
 * launches nothing) and must not be reused until the launch has finished on
 * `stream`. */
kleidicv_error_t kleidicv_hip_batch_workspace_size(const kleidicv_hip_op_desc *ops,
                                                   size_t count,
                                                   const kleidicv_hip_batch_tuning *tuning,
                                                   size_t *bytes, size_t *failed_index);
kleidicv_error_t kleidicv_hip_batch_launch(const kleidicv_hip_op_desc *ops, size_t count,
                                           const kleidicv_hip_batch_tuning *tuning,
                                           void *workspace, size_t workspace_bytes,
                                           void *stream, size_t *required_bytes,
                                           size_t *failed_index);

/* Waits for the device to finish outstanding work, then frees the batch. */
void kleidicv_hip_batch_destroy(kleidicv_hip_batch *batch);

/* The unbatched counterpart: each op through its single-image device entry
 * point (kleidicv_hip_*), in order, on `stream` (stages are kept by the
 * stream order).  Asynchronous.  On an error *failed_index (if not NULL)
 * names the op; the ops before it have been launched. */
kleidicv_error_t kleidicv_hip_run_ops(const kleidicv_hip_op_desc *ops, size_t count,
                                      void *stream, size_t *failed_index);

/* ------------------------------------------------------------------------
 * 5. Serving.  No counterpart in the reference: a continuous-batching
 *    request loop for one GPU.  Requests (an op descriptor over device
 *    memory, as for a batch) arrive at any time; the server admits each
 *    once its output fits the HBM budget, takes its output from a device
 *    ring, and launches it:
 *      - a large request (at least large_px output pixels) alone through
 *        its single-image entry point on slot 0, in arrival order; when more
 *        of them wait than slot 0 queues, up to large_batch of them share
 *        one batch launch;
 *      - small requests packed, up to max_batch_ops at a time, into one
 *        batch launch (kleidicv_hip_batch_launch) on the other slots (on
 *        slot 0 too when no large request waits); a lone small request runs
 *        alone.
 *    Each slot is a stream with up to `depth` launches queued; a request
 *    completes when its launch's event has completed.  Every request runs
 *    the kernels of its single-image entry point, so its output is the
 *    same alone and in any mix.
 *
 *    A server is driven by one host thread (not thread-safe).  Times are
 *    seconds on the caller's clock (kleidicv_hip_server_clock() for real
 *    time).
 * ------------------------------------------------------------------------ */

/*  hbm_budget: bytes of outputs and workspaces held at once (inputs are the
 *      caller's and resident already)
 *  workspace_reserve: work-list workspace of one slot, inside the budget
 *  large_px: output pixels from which a request is large
 *  slots: streams (>= 1); slot 0 takes the large requests
 *  max_batch_ops: small requests per batch launch (>= 1)
 *  depth: launches queued per slot (>= 1)
 *  large_batch: backlogged large requests per batch launch (1: always alone)
 *  release_done: 1 = an output returns to the ring when its request
 *      completes (the caller has copied or does not need it); 0 = it stays
 *      until kleidicv_hip_server_release()
 *  plain: 1 = the per-request baseline: every request is launched alone on
 *      slot 0 as soon as it is admitted (no batching, no queue limit)
 *  standin_polls: 0 = the GPU.  n > 0 = a host stand-in for tests without
 *      a GPU: nothing touches a device, pointers are never read, a batch
 *      needs 1024 workspace bytes per op, and a launch completes on the
 *      n-th poll of its event
 *  tuning: scheduling of the batch launches */
typedef struct {
  size_t hbm_budget;
  size_t workspace_reserve;
  size_t large_px;
  unsigned slots;
  unsigned max_batch_ops;
  unsigned depth;
  unsigned large_batch;
  unsigned release_done;
  unsigned plain;
  unsigned standin_polls;
  kleidicv_hip_batch_tuning tuning;
} kleidicv_hip_server_config;

typedef enum {
  KLEIDICV_HIP_REQUEST_QUEUED = 0,   /* arrived, waiting for budget */
  KLEIDICV_HIP_REQUEST_ADMITTED = 1, /* output held, waiting for a slot */
  KLEIDICV_HIP_REQUEST_LAUNCHED = 2,
  KLEIDICV_HIP_REQUEST_DONE = 3,
  KLEIDICV_HIP_REQUEST_FAILED = 4 /* `error` says why; nothing was written */
} kleidicv_hip_request_state;

typedef struct {
  uint64_t id;
  int state;     /* kleidicv_hip_request_state */
  int error;     /* kleidicv_error_t of a failed request */
  unsigned slot;
  unsigned alone;  /* 1: launched through its single-image entry point */
  uint64_t launch; /* launch number, from 1, in launch order */
  double arrival, admitted, launched, done;
  uint8_t *dst;
  size_t dst_stride;
  size_t out_bytes;
} kleidicv_hip_request_info;

typedef struct {
  uint64_t launches;
  size_t pending;         /* requests not yet completed */
  size_t out_in_use;      /* bytes of admitted outputs */
  size_t workspace_bytes; /* work-list workspaces of all slots */
  size_t peak;            /* most bytes of outputs and workspaces held at once */
  unsigned stalled;       /* 1: requests wait for ring space that only
                             kleidicv_hip_server_release() can return */
} kleidicv_hip_server_counters;

typedef struct kleidicv_hip_server kleidicv_hip_server;

/* ring/ring_bytes: device memory the outputs of requests without a `dst`
 * are taken from (rows of dst_width * channels bytes; NV12: 3 or 4
 * channels); may be NULL/0 when every request brings its own dst. */
kleidicv_error_t kleidicv_hip_server_create(const kleidicv_hip_server_config *config,
                                            void *ring, size_t ring_bytes,
                                            kleidicv_hip_server **server);

/* Queues `count` requests that arrived at arrival[i] (NULL: now is
 * unknown, 0); their ids are *first_id .. *first_id + count - 1.  A request
 * whose output can never fit the budget or the ring gives
 * KLEIDICV_ERROR_RANGE and nothing is queued. */
kleidicv_error_t kleidicv_hip_server_submit(kleidicv_hip_server *server,
                                            const kleidicv_hip_op_desc *ops, size_t count,
                                            const double *arrival, uint64_t *first_id);

/* One scheduling step at time `now`: retires completed launches, admits,
 * launches.  *pending (if not NULL) counts requests not yet completed. */
kleidicv_error_t kleidicv_hip_server_pump(kleidicv_hip_server *server, double now,
                                          size_t *pending);

/* Copies up to `capacity` completed (done or failed) requests, in
 * completion order, into `out` and forgets them; returns how many. */
size_t kleidicv_hip_server_take_completed(kleidicv_hip_server *server,
                                          kleidicv_hip_request_info *out, size_t capacity);

/* State of a request not yet taken. */
kleidicv_error_t kleidicv_hip_server_info(const kleidicv_hip_server *server, uint64_t id,
                                          kleidicv_hip_request_info *info);

/* Returns a completed request's output block to the ring. */
kleidicv_error_t kleidicv_hip_server_release(kleidicv_hip_server *server, uint64_t id);

kleidicv_error_t kleidicv_hip_server_stats(const kleidicv_hip_server *server,
                                           kleidicv_hip_server_counters *stats);

/* Replays a request stream in real time: request i is submitted once
 * offsets[i] seconds have passed since the call began (offsets ascending),
 * and the server is pumped until every request has completed.  results[i]
 * receives request i's final info; *t0 the clock at the start. */
kleidicv_error_t kleidicv_hip_server_replay(kleidicv_hip_server *server,
                                            const kleidicv_hip_op_desc *ops, size_t count,
                                            const double *offsets,
                                            kleidicv_hip_request_info *results, double *t0);

/* Monotonic clock in seconds. */
double kleidicv_hip_server_clock(void);

/* Waits for the server's launches to finish, then frees it. */
void kleidicv_hip_server_destroy(kleidicv_hip_server *server);

#ifdef __cplusplus
}
#endif

#endif /* KLEIDICV_HIP_KLEIDICV_H */
