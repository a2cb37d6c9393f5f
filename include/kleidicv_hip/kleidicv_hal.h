/*
 * OpenCV HAL entry points of kleidicv_hip: the hot-path members of the
 * reference's OpenCV adapter (adapters/opencv/kleidicv_hal.h:19-197), with
 * the same names, signatures and return codes, forwarding to this library's
 * GPU entry points.
 *
 * An OpenCV build patched for the reference adapter redirects its cv_hal_*
 * macros to kleidicv::hal::* (kleidicv_hal.h:199-702 wraps each call so a
 * non-OK return falls back to OpenCV's own code).  Linking that build against
 * libkleidicv_hip.so in place of the adapter's static libraries binds these
 * C++ symbols unchanged.  The extern "C" twins (kleidicv_hip_hal_*) take the
 * same arguments for callers that bind C symbols (ctypes, cgo, JNI).
 *
 * Return codes are OpenCV's: CV_HAL_ERROR_OK (0), CV_HAL_ERROR_NOT_IMPLEMENTED
 * (1, the caller falls back) and CV_HAL_ERROR_UNKNOWN (-1).  Depth, type,
 * border and interpolation arguments use OpenCV's encodings (CV_8U = 0,
 * type = depth + ((channels - 1) << 3); CV_HAL_BORDER_CONSTANT = 0,
 * REPLICATE = 1, REFLECT = 2, WRAP = 3, REFLECT_101 = 4, TRANSPARENT = 5,
 * ISOLATED = 16; CV_HAL_INTER_NEAREST = 0, LINEAR = 1, AREA = 3).
 *
 * Host and device pointers are both accepted (OpenCV passes host Mats; the
 * library stages them through device memory and returns when dst holds the
 * result).
 */
#ifndef KLEIDICV_HIP_KLEIDICV_HAL_H
#define KLEIDICV_HIP_KLEIDICV_HAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#include <cstdbool>

namespace kleidicv {
namespace hal {

using uchar = unsigned char;

/* kleidicv_hal.h:24, kleidicv_hal.cpp:191-200 */
int yuv422_to_bgr(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                  int width, int height, int dcn, bool swapBlue, int uIdx, int ycn);
/* kleidicv_hal.h:28, kleidicv_hal.cpp:180-189 (dcn is the source's channel
 * count) */
int bgr_to_yuv422(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                  int width, int height, int dcn, bool swapBlue, int uIdx, int ycn);
/* kleidicv_hal.h:39, kleidicv_hal.cpp:342-354 */
int bgr_to_yuv420_p(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                    int width, int height, int scn, bool swapBlue, int uIdx);
/* kleidicv_hal.h:43, kleidicv_hal.cpp:356-368 */
int bgr_to_yuv420_sp(const uchar *src_data, size_t src_step, uchar *y_data, size_t y_step,
                     uchar *uv_data, size_t uv_step, int width, int height, int scn,
                     bool swapBlue, int uIdx);
/* kleidicv_hal.h:64, kleidicv_hal.cpp:370-385 */
int bgr_to_yuv(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
               int width, int height, int depth, int scn, bool swapBlue, bool isCbCr);
/* kleidicv_hal.h:47, kleidicv_hal.cpp:286-293 */
int yuv_to_bgr_sp(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                  int dst_width, int dst_height, int dcn, bool swapBlue, int uIdx);
/* kleidicv_hal.h:51, kleidicv_hal.cpp:295-309 */
int yuv_to_bgr_sp_ex(const uchar *y_data, size_t y_step, const uchar *uv_data, size_t uv_step,
                     uchar *dst_data, size_t dst_step, int dst_width, int dst_height, int dcn,
                     bool swapBlue, int uIdx);
/* kleidicv_hal.h:56, kleidicv_hal.cpp:311-323 */
int yuv_to_bgr_p(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                 int dst_width, int dst_height, int dcn, bool swapBlue, int uIdx);
/* kleidicv_hal.h:60, kleidicv_hal.cpp:325-340 */
int yuv_to_bgr(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
               int width, int height, int depth, int dcn, bool swapBlue, bool isCbCr);
/* kleidicv_hal.h:72, kleidicv_hal.cpp:453-549 */
int separable_filter_2d(const uchar *src_data, size_t src_step, int src_type, uchar *dst_data,
                        size_t dst_step, int dst_type, int width, int height, int full_width,
                        int full_height, int offset_x, int offset_y, const uchar *kernelx_data,
                        int kernelx_len, const uchar *kernely_data, int kernely_len,
                        int kernel_type, int anchor_x, int anchor_y, double delta,
                        int borderType);
/* kleidicv_hal.h:80, kleidicv_hal.cpp:551-584 */
int gaussian_blur_binomial(const uchar *src_data, size_t src_step, uchar *dst_data,
                           size_t dst_step, int width, int height, int depth, int cn,
                           size_t margin_left, size_t margin_top, size_t margin_right,
                           size_t margin_bottom, size_t kernel_size, int border_type);
/* kleidicv_hal.h:87, kleidicv_hal.cpp:586-621 */
int gaussian_blur(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                  int width, int height, int depth, int cn, size_t margin_left,
                  size_t margin_top, size_t margin_right, size_t margin_bottom,
                  size_t kernel_width, size_t kernel_height, double sigma_x, double sigma_y,
                  int border_type);
/* kleidicv_hal.h:104, kleidicv_hal.cpp:735-819 */
int resize(int src_type, const uchar *src_data, size_t src_step, int src_width, int src_height,
           uchar *dst_data, size_t dst_step, int dst_width, int dst_height, double inv_scale_x,
           double inv_scale_y, int interpolation);
/* kleidicv_hal.h:115, kleidicv_hal.cpp:881-916 */
int medianBlur(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
               int width, int height, int depth, int cn, int ksize);
/* kleidicv_hal.h:129, kleidicv_hal.cpp:1562-1592 */
int pyrdown(const uchar *src_data, size_t src_step, int src_width, int src_height,
            uchar *dst_data, size_t dst_step, int dst_width, int dst_height, int depth, int cn,
            int border_type);
/* kleidicv_hal.h:175, kleidicv_hal.cpp:1458-1519 */
int remap_s16(int src_type, const uchar *src_data, size_t src_step, int src_width,
              int src_height, uchar *dst_data, size_t dst_step, int dst_width, int dst_height,
              int16_t *mapx, size_t mapx_step, uint16_t *mapy, size_t mapy_step,
              int interpolation, int border_type, const double border_value_f64[4]);
/* kleidicv_hal.h:181, kleidicv_hal.cpp:1521-1560 */
int remap_f32(int src_type, const uchar *src_data, size_t src_step, int src_width,
              int src_height, uchar *dst_data, size_t dst_step, int dst_width, int dst_height,
              const float *mapx, size_t mapx_step, const float *mapy, size_t mapy_step,
              int interpolation, int border_type, const double border_value[4]);
/* kleidicv_hal.h:187, kleidicv_hal.cpp:1608-1641 */
int warp_perspective(int src_type, const uchar *src_data, size_t src_step, int src_width,
                     int src_height, uchar *dst_data, size_t dst_step, int dst_width,
                     int dst_height, const double transformation[9], int interpolation,
                     int borderType, const double borderValue[4]);

}  // namespace hal
}  // namespace kleidicv

extern "C" {
#else
#include <stdbool.h>
#endif

/* C twins of the functions above (same arguments, same return codes). */
int kleidicv_hip_hal_bgr_to_yuv422(const unsigned char *src_data, size_t src_step,
                                   unsigned char *dst_data, size_t dst_step, int width,
                                   int height, int dcn, bool swapBlue, int uIdx, int ycn);
int kleidicv_hip_hal_bgr_to_yuv420_p(const unsigned char *src_data, size_t src_step,
                                     unsigned char *dst_data, size_t dst_step, int width,
                                     int height, int scn, bool swapBlue, int uIdx);
int kleidicv_hip_hal_bgr_to_yuv420_sp(const unsigned char *src_data, size_t src_step,
                                      unsigned char *y_data, size_t y_step,
                                      unsigned char *uv_data, size_t uv_step, int width,
                                      int height, int scn, bool swapBlue, int uIdx);
int kleidicv_hip_hal_bgr_to_yuv(const unsigned char *src_data, size_t src_step,
                                unsigned char *dst_data, size_t dst_step, int width,
                                int height, int depth, int scn, bool swapBlue, bool isCbCr);
int kleidicv_hip_hal_yuv422_to_bgr(const unsigned char *src_data, size_t src_step,
                                   unsigned char *dst_data, size_t dst_step, int width,
                                   int height, int dcn, bool swapBlue, int uIdx, int ycn);
int kleidicv_hip_hal_yuv_to_bgr_sp(const unsigned char *src_data, size_t src_step,
                                   unsigned char *dst_data, size_t dst_step, int dst_width,
                                   int dst_height, int dcn, bool swapBlue, int uIdx);
int kleidicv_hip_hal_yuv_to_bgr_sp_ex(const unsigned char *y_data, size_t y_step,
                                      const unsigned char *uv_data, size_t uv_step,
                                      unsigned char *dst_data, size_t dst_step, int dst_width,
                                      int dst_height, int dcn, bool swapBlue, int uIdx);
int kleidicv_hip_hal_yuv_to_bgr_p(const unsigned char *src_data, size_t src_step,
                                  unsigned char *dst_data, size_t dst_step, int dst_width,
                                  int dst_height, int dcn, bool swapBlue, int uIdx);
int kleidicv_hip_hal_yuv_to_bgr(const unsigned char *src_data, size_t src_step,
                                unsigned char *dst_data, size_t dst_step, int width,
                                int height, int depth, int dcn, bool swapBlue, bool isCbCr);
int kleidicv_hip_hal_separable_filter_2d(
    const unsigned char *src_data, size_t src_step, int src_type, unsigned char *dst_data,
    size_t dst_step, int dst_type, int width, int height, int full_width, int full_height,
    int offset_x, int offset_y, const unsigned char *kernelx_data, int kernelx_len,
    const unsigned char *kernely_data, int kernely_len, int kernel_type, int anchor_x,
    int anchor_y, double delta, int borderType);
int kleidicv_hip_hal_gaussian_blur_binomial(const unsigned char *src_data, size_t src_step,
                                            unsigned char *dst_data, size_t dst_step,
                                            int width, int height, int depth, int cn,
                                            size_t margin_left, size_t margin_top,
                                            size_t margin_right, size_t margin_bottom,
                                            size_t kernel_size, int border_type);
int kleidicv_hip_hal_gaussian_blur(const unsigned char *src_data, size_t src_step,
                                   unsigned char *dst_data, size_t dst_step, int width,
                                   int height, int depth, int cn, size_t margin_left,
                                   size_t margin_top, size_t margin_right,
                                   size_t margin_bottom, size_t kernel_width,
                                   size_t kernel_height, double sigma_x, double sigma_y,
                                   int border_type);
int kleidicv_hip_hal_resize(int src_type, const unsigned char *src_data, size_t src_step,
                            int src_width, int src_height, unsigned char *dst_data,
                            size_t dst_step, int dst_width, int dst_height,
                            double inv_scale_x, double inv_scale_y, int interpolation);
int kleidicv_hip_hal_medianBlur(const unsigned char *src_data, size_t src_step,
                                unsigned char *dst_data, size_t dst_step, int width,
                                int height, int depth, int cn, int ksize);
int kleidicv_hip_hal_pyrdown(const unsigned char *src_data, size_t src_step, int src_width,
                             int src_height, unsigned char *dst_data, size_t dst_step,
                             int dst_width, int dst_height, int depth, int cn,
                             int border_type);
int kleidicv_hip_hal_remap_s16(int src_type, const unsigned char *src_data, size_t src_step,
                               int src_width, int src_height, unsigned char *dst_data,
                               size_t dst_step, int dst_width, int dst_height, int16_t *mapx,
                               size_t mapx_step, uint16_t *mapy, size_t mapy_step,
                               int interpolation, int border_type,
                               const double border_value_f64[4]);
int kleidicv_hip_hal_remap_f32(int src_type, const unsigned char *src_data, size_t src_step,
                               int src_width, int src_height, unsigned char *dst_data,
                               size_t dst_step, int dst_width, int dst_height,
                               const float *mapx, size_t mapx_step, const float *mapy,
                               size_t mapy_step, int interpolation, int border_type,
                               const double border_value[4]);
int kleidicv_hip_hal_warp_perspective(int src_type, const unsigned char *src_data,
                                      size_t src_step, int src_width, int src_height,
                                      unsigned char *dst_data, size_t dst_step, int dst_width,
                                      int dst_height, const double transformation[9],
                                      int interpolation, int borderType,
                                      const double borderValue[4]);

#ifdef __cplusplus
}
#endif

#endif /* KLEIDICV_HIP_KLEIDICV_HAL_H */
