// OpenCV HAL entry points (kleidicv_hal.h in include/kleidicv_hip): the
// reference adapter's policy for the hot-path operations, forwarding to this
// library's drop-in entry points.
//
// Policy follows adapters/opencv/kleidicv_hal.cpp line by line where it
// decides what OpenCV hands over (ROI margins, in-place calls, depths,
// border and interpolation maps, error mapping).  Two groups of its checks
// are not carried over, because they describe the Neon library rather than
// the operation:
//   - the resize declines for single-threaded OpenCV builds
//     (kleidicv_hal.cpp:776-787), which exist because Neon lost to OpenCV's
//     own code there; the GPU kernels do not;
//   - the thread-count heuristics (MULTITHREAD_MIN_ELEMENTS_*), since one
//     GPU launch replaces the stripe scheduler at every size.
// Channel counts the reference library declines are declined here too
// (warp: CV_8UC1 only; remap f32: 1 or 2 channels), so a patched OpenCV
// behaves as with the reference adapter.  Depths other than 8-bit unsigned
// return NOT_IMPLEMENTED (OpenCV falls back to its own code).

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "kleidicv_hip/kleidicv.h"
#include "kleidicv_hip/kleidicv_hal.h"

namespace {

// OpenCV's encodings (opencv2/core/hal/interface.h, imgproc/hal/interface.h).
constexpr int kHalOk = 0;
constexpr int kHalNotImplemented = 1;
constexpr int kHalUnknown = -1;
constexpr int kCv8U = 0;
constexpr int kCvCnShift = 3;
constexpr int kBorderConstant = 0, kBorderReplicate = 1, kBorderReflect = 2, kBorderWrap = 3,
              kBorderReflect101 = 4, kBorderTransparent = 5, kBorderIsolated = 16;
constexpr int kInterNearest = 0, kInterLinear = 1, kInterArea = 3;

int depth_of(int type) { return type & ((1 << kCvCnShift) - 1); }
int channels_of(int type) { return ((type >> kCvCnShift) & 511) + 1; }

// kleidicv_hal.cpp:78-92
int convert_error(kleidicv_error_t e) {
  switch (e) {
    case KLEIDICV_OK:
      return kHalOk;
    case KLEIDICV_ERROR_NOT_IMPLEMENTED:
    case KLEIDICV_ERROR_ALIGNMENT:
      return kHalNotImplemented;
    default:
      return kHalUnknown;
  }
}

// kleidicv_hal.cpp:405-435
bool border_from_opencv(int cv, kleidicv_border_type_t *out) {
  switch (cv) {
    case kBorderConstant: *out = KLEIDICV_BORDER_TYPE_CONSTANT; return true;
    case kBorderReplicate: *out = KLEIDICV_BORDER_TYPE_REPLICATE; return true;
    case kBorderReflect: *out = KLEIDICV_BORDER_TYPE_REFLECT; return true;
    case kBorderReflect101: *out = KLEIDICV_BORDER_TYPE_REVERSE; return true;
    case kBorderWrap: *out = KLEIDICV_BORDER_TYPE_WRAP; return true;
    case kBorderTransparent: *out = KLEIDICV_BORDER_TYPE_TRANSPARENT; return true;
    case kBorderIsolated: *out = KLEIDICV_BORDER_TYPE_NONE; return true;
    default: return false;
  }
}

// kleidicv_hal.cpp:437-451
bool interpolation_from_opencv(int cv, kleidicv_interpolation_type_t *out) {
  switch (cv) {
    case kInterNearest: *out = KLEIDICV_INTERPOLATION_NEAREST; return true;
    case kInterLinear: *out = KLEIDICV_INTERPOLATION_LINEAR; return true;
    default: return false;
  }
}

// cv::saturate_cast<uint8_t>(double): round half to even, then clamp
// (kleidicv_hal.cpp:109-117).
struct BorderValue {
  uint8_t v[KLEIDICV_MAXIMUM_CHANNEL_COUNT];
};

BorderValue border_value_u8(const double f64[4]) {
  BorderValue b{};
  for (int i = 0; i < 4; ++i) {
    const double r = f64 ? std::nearbyint(f64[i]) : 0.0;
    b.v[i] = static_cast<uint8_t>(r < 0.0 ? 0.0 : (r > 255.0 ? 255.0 : r));
  }
  return b;
}

// kleidicv_hal.cpp:58-76
kleidicv_color_conversion_t color_type(int channels, bool bgr, bool vu, bool uyvy,
                                       kleidicv_color_conversion_t base) {
  unsigned v = static_cast<unsigned>(base);
  if (vu) v |= KLEIDICV_COLOR_CONVERSION_FLAG_VU;
  if (bgr) v |= KLEIDICV_COLOR_CONVERSION_FLAG_BGR;
  if (channels == 4) v |= KLEIDICV_COLOR_CONVERSION_FLAG_ALPHA;
  if (uyvy) v |= KLEIDICV_COLOR_CONVERSION_FLAG_CHROMA_FIRST;
  return static_cast<kleidicv_color_conversion_t>(v);
}

}  // namespace

namespace kleidicv {
namespace hal {

int yuv422_to_bgr(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                  int width, int height, int dcn, bool swapBlue, int uIdx, int ycn) {
  return convert_error(kleidicv_yuv_to_rgb_u8(
      src_data, src_step, dst_data, dst_step, static_cast<size_t>(width),
      static_cast<size_t>(height),
      color_type(dcn, !swapBlue, uIdx != 0, ycn != 0, KLEIDICV_COLOR_CONVERSION_FMT_YUV422)));
}

int bgr_to_yuv422(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                  int width, int height, int dcn, bool swapBlue, int uIdx, int ycn) {
  return convert_error(kleidicv_rgb_to_yuv_u8(
      src_data, src_step, dst_data, dst_step, static_cast<size_t>(width),
      static_cast<size_t>(height),
      color_type(dcn, !swapBlue, uIdx != 0, ycn != 0, KLEIDICV_COLOR_CONVERSION_FMT_YUV422)));
}

int bgr_to_yuv420_p(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                    int width, int height, int scn, bool swapBlue, int uIdx) {
  // OpenCV's uIdx 2 is YV12 here (kleidicv_hal.cpp:345)
  return convert_error(kleidicv_rgb_to_yuv_u8(
      src_data, src_step, dst_data, dst_step, static_cast<size_t>(width),
      static_cast<size_t>(height),
      color_type(scn, !swapBlue, uIdx == 2, false, KLEIDICV_COLOR_CONVERSION_FMT_YUV420P)));
}

int bgr_to_yuv420_sp(const uchar *src_data, size_t src_step, uchar *y_data, size_t y_step,
                     uchar *uv_data, size_t uv_step, int width, int height, int scn,
                     bool swapBlue, int uIdx) {
  return convert_error(kleidicv_rgb_to_yuv_semiplanar_u8(
      src_data, src_step, y_data, y_step, uv_data, uv_step, static_cast<size_t>(width),
      static_cast<size_t>(height),
      color_type(scn, !swapBlue, uIdx == 2, false, KLEIDICV_COLOR_CONVERSION_FMT_YUV420SP)));
}

int bgr_to_yuv(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
               int width, int height, int depth, int scn, bool swapBlue, bool isCbCr) {
  if (depth != kCv8U || isCbCr) return kHalNotImplemented;
  return convert_error(kleidicv_rgb_to_yuv_u8(
      src_data, src_step, dst_data, dst_step, static_cast<size_t>(width),
      static_cast<size_t>(height),
      color_type(scn, !swapBlue, false, false, KLEIDICV_COLOR_CONVERSION_FMT_YUV444)));
}

int yuv_to_bgr_sp_ex(const uchar *y_data, size_t y_step, const uchar *uv_data, size_t uv_step,
                     uchar *dst_data, size_t dst_step, int dst_width, int dst_height, int dcn,
                     bool swapBlue, int uIdx) {
  return convert_error(kleidicv_yuv_semiplanar_to_rgb_u8(
      y_data, y_step, uv_data, uv_step, dst_data, dst_step, static_cast<size_t>(dst_width),
      static_cast<size_t>(dst_height),
      color_type(dcn, !swapBlue, uIdx != 0, false, KLEIDICV_COLOR_CONVERSION_FMT_YUV420SP)));
}

int yuv_to_bgr_sp(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                  int dst_width, int dst_height, int dcn, bool swapBlue, int uIdx) {
  // the chroma plane follows the luma plane (kleidicv_hal.cpp:286-293)
  const uchar *uv_data = src_data + static_cast<size_t>(dst_height) * src_step;
  return yuv_to_bgr_sp_ex(src_data, src_step, uv_data, src_step, dst_data, dst_step,
                          dst_width, dst_height, dcn, swapBlue, uIdx);
}

int yuv_to_bgr_p(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                 int dst_width, int dst_height, int dcn, bool swapBlue, int uIdx) {
  return convert_error(kleidicv_yuv_to_rgb_u8(
      src_data, src_step, dst_data, dst_step, static_cast<size_t>(dst_width),
      static_cast<size_t>(dst_height),
      color_type(dcn, !swapBlue, uIdx != 0, false, KLEIDICV_COLOR_CONVERSION_FMT_YUV420P)));
}

int yuv_to_bgr(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
               int width, int height, int depth, int dcn, bool swapBlue, bool isCbCr) {
  if (depth != kCv8U || isCbCr) return kHalNotImplemented;
  return convert_error(kleidicv_yuv_to_rgb_u8(
      src_data, src_step, dst_data, dst_step, static_cast<size_t>(width),
      static_cast<size_t>(height),
      color_type(dcn, !swapBlue, false, false, KLEIDICV_COLOR_CONVERSION_FMT_YUV444)));
}

int separable_filter_2d(const uchar *src_data, size_t src_step, int src_type, uchar *dst_data,
                        size_t dst_step, int dst_type, int width, int height, int full_width,
                        int full_height, int offset_x, int offset_y, const uchar *kernelx_data,
                        int kernelx_len, const uchar *kernely_data, int kernely_len,
                        int kernel_type, int anchor_x, int anchor_y, double delta,
                        int borderType) {
  // a submatrix without BORDER_ISOLATED would read outside the ROI
  if (offset_x != 0 || offset_y != 0 || full_width - width - offset_x != 0 ||
      full_height - height - offset_y != 0)
    return kHalNotImplemented;
  if (src_data == dst_data) return kHalNotImplemented;
  if (src_type != dst_type) return kHalNotImplemented;
  if (depth_of(src_type) != depth_of(kernel_type)) return kHalNotImplemented;
  if (channels_of(kernel_type) != 1) return kHalNotImplemented;
  if (depth_of(src_type) != kCv8U) return kHalNotImplemented;  // 16U: not on this path
  kleidicv_border_type_t border;
  if (!border_from_opencv(borderType, &border)) return kHalNotImplemented;
  if (border != KLEIDICV_BORDER_TYPE_REPLICATE && border != KLEIDICV_BORDER_TYPE_REFLECT &&
      border != KLEIDICV_BORDER_TYPE_WRAP && border != KLEIDICV_BORDER_TYPE_REVERSE)
    return kHalNotImplemented;
  if (anchor_x > -1 && anchor_x != (kernelx_len >> 1)) return kHalNotImplemented;
  if (anchor_y > -1 && anchor_y != (kernely_len >> 1)) return kHalNotImplemented;
  if (delta != 0.0) return kHalNotImplemented;
  return convert_error(kleidicv_separable_filter_2d_u8(
      src_data, src_step, dst_data, dst_step, static_cast<size_t>(width),
      static_cast<size_t>(height), static_cast<size_t>(channels_of(src_type)), kernelx_data,
      static_cast<size_t>(kernelx_len), kernely_data, static_cast<size_t>(kernely_len),
      border));
}

int gaussian_blur_binomial(const uchar *src_data, size_t src_step, uchar *dst_data,
                           size_t dst_step, int width, int height, int depth, int cn,
                           size_t margin_left, size_t margin_top, size_t margin_right,
                           size_t margin_bottom, size_t kernel_size, int border_type) {
  if (src_data == dst_data) return kHalNotImplemented;
  if (margin_left != 0 || margin_top != 0 || margin_right != 0 || margin_bottom != 0)
    return kHalNotImplemented;
  if (depth != kCv8U) return kHalNotImplemented;
  kleidicv_border_type_t border;
  if (!border_from_opencv(border_type, &border)) return kHalNotImplemented;
  return convert_error(kleidicv_gaussian_blur_u8(
      src_data, src_step, dst_data, dst_step, static_cast<size_t>(width),
      static_cast<size_t>(height), static_cast<size_t>(cn), kernel_size, kernel_size, 0.0f,
      0.0f, border));
}

int gaussian_blur(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
                  int width, int height, int depth, int cn, size_t margin_left,
                  size_t margin_top, size_t margin_right, size_t margin_bottom,
                  size_t kernel_width, size_t kernel_height, double sigma_x, double sigma_y,
                  int border_type) {
  if (src_data == dst_data) return kHalNotImplemented;
  if (margin_left != 0 || margin_top != 0 || margin_right != 0 || margin_bottom != 0)
    return kHalNotImplemented;
  if (depth != kCv8U) return kHalNotImplemented;
  kleidicv_border_type_t border;
  if (!border_from_opencv(border_type, &border)) return kHalNotImplemented;
  return convert_error(kleidicv_gaussian_blur_u8(
      src_data, src_step, dst_data, dst_step, static_cast<size_t>(width),
      static_cast<size_t>(height), static_cast<size_t>(cn), kernel_width, kernel_height,
      static_cast<float>(sigma_x), static_cast<float>(sigma_y), border));
}

int resize(int src_type, const uchar *src_data, size_t src_step, int src_width, int src_height,
           uchar *dst_data, size_t dst_step, int dst_width, int dst_height, double inv_scale_x,
           double inv_scale_y, int interpolation) {
  if (src_data == dst_data) return kHalNotImplemented;
  const int channels = channels_of(src_type);
  if (depth_of(src_type) != kCv8U) return kHalNotImplemented;  // f32: not on this path
  auto call = [&] {
    return convert_error(kleidicv_resize_linear_u8(
        src_data, src_step, static_cast<size_t>(src_width), static_cast<size_t>(src_height),
        dst_data, dst_step, static_cast<size_t>(dst_width), static_cast<size_t>(dst_height),
        static_cast<size_t>(channels)));
  };
  // 0.5 x 0.5: INTER_LINEAR and INTER_AREA agree (kleidicv_hal.cpp:747-761)
  if (inv_scale_x == 0.5 && inv_scale_y == 0.5 && channels >= 1 && channels <= 4 &&
      (interpolation == kInterLinear || interpolation == kInterArea))
    return call();
  if (interpolation != kInterLinear) return kHalNotImplemented;
  if ((inv_scale_x != 0 && std::fabs(inv_scale_x * src_width - dst_width) > FLT_EPSILON) ||
      (inv_scale_y != 0 && std::fabs(inv_scale_y * src_height - dst_height) > FLT_EPSILON))
    return kHalNotImplemented;
  return call();
}

int medianBlur(const uchar *src_data, size_t src_step, uchar *dst_data, size_t dst_step,
               int width, int height, int depth, int cn, int ksize) {
  if (src_data == dst_data) return kHalNotImplemented;
  if (depth != kCv8U) return kHalNotImplemented;  // s16/u16/f32: not on this path
  // OpenCV's medianBlur replicates the border (kleidicv_hal.cpp:890-895)
  return convert_error(kleidicv_median_blur_u8(
      src_data, src_step, dst_data, dst_step, static_cast<size_t>(width),
      static_cast<size_t>(height), static_cast<size_t>(cn), static_cast<size_t>(ksize),
      static_cast<size_t>(ksize), KLEIDICV_BORDER_TYPE_REPLICATE));
}

int pyrdown(const uchar *src_data, size_t src_step, int src_width, int src_height,
            uchar *dst_data, size_t dst_step, int dst_width, int dst_height, int depth, int cn,
            int border_type) {
  if (src_data == dst_data) return kHalNotImplemented;
  if (depth != kCv8U) return kHalNotImplemented;
  if (dst_width != (src_width + 1) / 2 || dst_height != (src_height + 1) / 2)
    return kHalNotImplemented;
  kleidicv_border_type_t border;
  if (!border_from_opencv(border_type, &border)) return kHalNotImplemented;
  return convert_error(kleidicv_blur_and_downsample_u8(
      src_data, src_step, static_cast<size_t>(src_width), static_cast<size_t>(src_height),
      dst_data, dst_step, static_cast<size_t>(cn), border));
}

int remap_s16(int src_type, const uchar *src_data, size_t src_step, int src_width,
              int src_height, uchar *dst_data, size_t dst_step, int dst_width, int dst_height,
              int16_t *mapx, size_t mapx_step, uint16_t *mapy, size_t mapy_step,
              int interpolation, int border_type, const double border_value_f64[4]) {
  kleidicv_border_type_t border;
  if (!border_from_opencv(border_type, &border)) return kHalNotImplemented;
  // INTER_NEAREST (kleidicv_remap_s16_u8) and 16-bit images are not on this
  // path; OpenCV's mapx is the interleaved integer part, mapy the fractions.
  if (interpolation != kInterLinear || depth_of(src_type) != kCv8U) return kHalNotImplemented;
  const BorderValue bv = border_value_u8(border_value_f64);
  return convert_error(kleidicv_remap_s16point5_u8(
      src_data, src_step, static_cast<size_t>(src_width), static_cast<size_t>(src_height),
      dst_data, dst_step, static_cast<size_t>(dst_width), static_cast<size_t>(dst_height),
      static_cast<size_t>(channels_of(src_type)), mapx, mapx_step, mapy, mapy_step, border,
      bv.v));
}

int remap_f32(int src_type, const uchar *src_data, size_t src_step, int src_width,
              int src_height, uchar *dst_data, size_t dst_step, int dst_width, int dst_height,
              const float *mapx, size_t mapx_step, const float *mapy, size_t mapy_step,
              int interpolation, int border_type, const double border_value[4]) {
  kleidicv_border_type_t border;
  if (!border_from_opencv(border_type, &border)) return kHalNotImplemented;
  kleidicv_interpolation_type_t interp;
  if (!interpolation_from_opencv(interpolation, &interp)) return kHalNotImplemented;
  if (depth_of(src_type) != kCv8U) return kHalNotImplemented;
  const int channels = channels_of(src_type);
  if (channels > 2) return kHalNotImplemented;  // the reference library's limit (remap.h:55-74)
  const BorderValue bv = border_value_u8(border_value);
  return convert_error(kleidicv_remap_f32_u8(
      src_data, src_step, static_cast<size_t>(src_width), static_cast<size_t>(src_height),
      dst_data, dst_step, static_cast<size_t>(dst_width), static_cast<size_t>(dst_height),
      static_cast<size_t>(channels), mapx, mapx_step, mapy, mapy_step, interp, border, bv.v));
}

int warp_perspective(int src_type, const uchar *src_data, size_t src_step, int src_width,
                     int src_height, uchar *dst_data, size_t dst_step, int dst_width,
                     int dst_height, const double transformation[9], int interpolation,
                     int borderType, const double borderValue[4]) {
  kleidicv_border_type_t border;
  if (!border_from_opencv(borderType, &border)) return kHalNotImplemented;
  float t[9];
  for (int i = 0; i < 9; ++i) t[i] = static_cast<float>(transformation[i]);
  kleidicv_interpolation_type_t interp;
  if (!interpolation_from_opencv(interpolation, &interp)) return kHalNotImplemented;
  const BorderValue bv = border_value_u8(borderValue);
  if (src_type != kCv8U) return kHalNotImplemented;  // CV_8UC1 only, as the reference
  return convert_error(kleidicv_warp_perspective_u8(
      src_data, src_step, static_cast<size_t>(src_width), static_cast<size_t>(src_height),
      dst_data, dst_step, static_cast<size_t>(dst_width), static_cast<size_t>(dst_height), t,
      1, interp, border, bv.v));
}

}  // namespace hal
}  // namespace kleidicv

extern "C" {

namespace H = kleidicv::hal;

int kleidicv_hip_hal_bgr_to_yuv422(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *dst_data,
    size_t dst_step,
    int width,
    int height,
    int dcn,
    bool swapBlue,
    int uIdx,
    int ycn) {
  return H::bgr_to_yuv422(src_data, src_step, dst_data, dst_step, width, height, dcn, swapBlue, uIdx, ycn);
}

int kleidicv_hip_hal_bgr_to_yuv420_p(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *dst_data,
    size_t dst_step,
    int width,
    int height,
    int scn,
    bool swapBlue,
    int uIdx) {
  return H::bgr_to_yuv420_p(src_data, src_step, dst_data, dst_step, width, height, scn, swapBlue, uIdx);
}

int kleidicv_hip_hal_bgr_to_yuv420_sp(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *y_data,
    size_t y_step,
    unsigned char *uv_data,
    size_t uv_step,
    int width,
    int height,
    int scn,
    bool swapBlue,
    int uIdx) {
  return H::bgr_to_yuv420_sp(src_data, src_step, y_data, y_step, uv_data, uv_step, width, height, scn, swapBlue, uIdx);
}

int kleidicv_hip_hal_bgr_to_yuv(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *dst_data,
    size_t dst_step,
    int width,
    int height,
    int depth,
    int scn,
    bool swapBlue,
    bool isCbCr) {
  return H::bgr_to_yuv(src_data, src_step, dst_data, dst_step, width, height, depth, scn, swapBlue, isCbCr);
}

int kleidicv_hip_hal_yuv422_to_bgr(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *dst_data,
    size_t dst_step,
    int width,
    int height,
    int dcn,
    bool swapBlue,
    int uIdx,
    int ycn) {
  return H::yuv422_to_bgr(src_data, src_step, dst_data, dst_step, width, height, dcn, swapBlue, uIdx, ycn);
}

int kleidicv_hip_hal_yuv_to_bgr_sp(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *dst_data,
    size_t dst_step,
    int dst_width,
    int dst_height,
    int dcn,
    bool swapBlue,
    int uIdx) {
  return H::yuv_to_bgr_sp(src_data, src_step, dst_data, dst_step, dst_width, dst_height, dcn, swapBlue, uIdx);
}

int kleidicv_hip_hal_yuv_to_bgr_sp_ex(
    const unsigned char *y_data,
    size_t y_step,
    const unsigned char *uv_data,
    size_t uv_step,
    unsigned char *dst_data,
    size_t dst_step,
    int dst_width,
    int dst_height,
    int dcn,
    bool swapBlue,
    int uIdx) {
  return H::yuv_to_bgr_sp_ex(y_data, y_step, uv_data, uv_step, dst_data, dst_step, dst_width, dst_height, dcn, swapBlue, uIdx);
}

int kleidicv_hip_hal_yuv_to_bgr_p(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *dst_data,
    size_t dst_step,
    int dst_width,
    int dst_height,
    int dcn,
    bool swapBlue,
    int uIdx) {
  return H::yuv_to_bgr_p(src_data, src_step, dst_data, dst_step, dst_width, dst_height, dcn, swapBlue, uIdx);
}

int kleidicv_hip_hal_yuv_to_bgr(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *dst_data,
    size_t dst_step,
    int width,
    int height,
    int depth,
    int dcn,
    bool swapBlue,
    bool isCbCr) {
  return H::yuv_to_bgr(src_data, src_step, dst_data, dst_step, width, height, depth, dcn, swapBlue, isCbCr);
}

int kleidicv_hip_hal_separable_filter_2d(
    const unsigned char *src_data,
    size_t src_step,
    int src_type,
    unsigned char *dst_data,
    size_t dst_step,
    int dst_type,
    int width,
    int height,
    int full_width,
    int full_height,
    int offset_x,
    int offset_y,
    const unsigned char *kernelx_data,
    int kernelx_len,
    const unsigned char *kernely_data,
    int kernely_len,
    int kernel_type,
    int anchor_x,
    int anchor_y,
    double delta,
    int borderType) {
  return H::separable_filter_2d(src_data, src_step, src_type, dst_data, dst_step, dst_type, width, height, full_width, full_height, offset_x, offset_y, kernelx_data, kernelx_len, kernely_data, kernely_len, kernel_type, anchor_x, anchor_y, delta, borderType);
}

int kleidicv_hip_hal_gaussian_blur_binomial(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *dst_data,
    size_t dst_step,
    int width,
    int height,
    int depth,
    int cn,
    size_t margin_left,
    size_t margin_top,
    size_t margin_right,
    size_t margin_bottom,
    size_t kernel_size,
    int border_type) {
  return H::gaussian_blur_binomial(src_data, src_step, dst_data, dst_step, width, height, depth, cn, margin_left, margin_top, margin_right, margin_bottom, kernel_size, border_type);
}

int kleidicv_hip_hal_gaussian_blur(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *dst_data,
    size_t dst_step,
    int width,
    int height,
    int depth,
    int cn,
    size_t margin_left,
    size_t margin_top,
    size_t margin_right,
    size_t margin_bottom,
    size_t kernel_width,
    size_t kernel_height,
    double sigma_x,
    double sigma_y,
    int border_type) {
  return H::gaussian_blur(src_data, src_step, dst_data, dst_step, width, height, depth, cn, margin_left, margin_top, margin_right, margin_bottom, kernel_width, kernel_height, sigma_x, sigma_y, border_type);
}

int kleidicv_hip_hal_resize(
    int src_type,
    const unsigned char *src_data,
    size_t src_step,
    int src_width,
    int src_height,
    unsigned char *dst_data,
    size_t dst_step,
    int dst_width,
    int dst_height,
    double inv_scale_x,
    double inv_scale_y,
    int interpolation) {
  return H::resize(src_type, src_data, src_step, src_width, src_height, dst_data, dst_step, dst_width, dst_height, inv_scale_x, inv_scale_y, interpolation);
}

int kleidicv_hip_hal_medianBlur(
    const unsigned char *src_data,
    size_t src_step,
    unsigned char *dst_data,
    size_t dst_step,
    int width,
    int height,
    int depth,
    int cn,
    int ksize) {
  return H::medianBlur(src_data, src_step, dst_data, dst_step, width, height, depth, cn, ksize);
}

int kleidicv_hip_hal_pyrdown(
    const unsigned char *src_data,
    size_t src_step,
    int src_width,
    int src_height,
    unsigned char *dst_data,
    size_t dst_step,
    int dst_width,
    int dst_height,
    int depth,
    int cn,
    int border_type) {
  return H::pyrdown(src_data, src_step, src_width, src_height, dst_data, dst_step, dst_width, dst_height, depth, cn, border_type);
}

int kleidicv_hip_hal_remap_s16(
    int src_type,
    const unsigned char *src_data,
    size_t src_step,
    int src_width,
    int src_height,
    unsigned char *dst_data,
    size_t dst_step,
    int dst_width,
    int dst_height,
    int16_t *mapx,
    size_t mapx_step,
    uint16_t *mapy,
    size_t mapy_step,
    int interpolation,
    int border_type,
    const double border_value_f64[4]) {
  return H::remap_s16(src_type, src_data, src_step, src_width, src_height, dst_data, dst_step, dst_width, dst_height, mapx, mapx_step, mapy, mapy_step, interpolation, border_type, border_value_f64);
}

int kleidicv_hip_hal_remap_f32(
    int src_type,
    const unsigned char *src_data,
    size_t src_step,
    int src_width,
    int src_height,
    unsigned char *dst_data,
    size_t dst_step,
    int dst_width,
    int dst_height,
    const float *mapx,
    size_t mapx_step,
    const float *mapy,
    size_t mapy_step,
    int interpolation,
    int border_type,
    const double border_value[4]) {
  return H::remap_f32(src_type, src_data, src_step, src_width, src_height, dst_data, dst_step, dst_width, dst_height, mapx, mapx_step, mapy, mapy_step, interpolation, border_type, border_value);
}

int kleidicv_hip_hal_warp_perspective(
    int src_type,
    const unsigned char *src_data,
    size_t src_step,
    int src_width,
    int src_height,
    unsigned char *dst_data,
    size_t dst_step,
    int dst_width,
    int dst_height,
    const double transformation[9],
    int interpolation,
    int borderType,
    const double borderValue[4]) {
  return H::warp_perspective(src_type, src_data, src_step, src_width, src_height, dst_data, dst_step, dst_width, dst_height, transformation, interpolation, borderType, borderValue);
}

}  // extern "C"
