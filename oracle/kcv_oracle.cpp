// CPU oracle for the KleidiCV u8 image hot path.
//
// TEST INFRASTRUCTURE ONLY: a scalar restatement of the reference algorithms,
// used as the checker by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg.  The product never links or calls it.
//
// Built with -ffp-contract=off: wherever the reference's clang build fuses a
// multiply-add (vmlaq_f32 / scalar a*b+c under the default -ffp-contract=on),
// this file calls std::fmaf explicitly, and nowhere else.
//
// Parity status of each function is listed in the header comment of the
// function.  "Pinned" names the independent check in tests/ that fixes it.

#include "kcv_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

enum {
  kBorderConstant = 0,
  kBorderReplicate = 1,
  kBorderReflect = 2,
  kBorderWrap = 3,
  kBorderReverse = 4,
};

enum { kErrOk = 0, kErrNotImplemented = 1, kErrNull = 2, kErrRange = 3 };

// Single-reflection border mapping; valid while |overhang| <= limit, which the
// reference guarantees through its w,h >= k-1 precondition.
// Reference: src/filters/median_blur_border_handling.h:14-56, and the offset
// tables of include/kleidicv/workspace/border_{3x3,5x5,7x7}.h which encode
// the same mapping row by row.
long border_index(long i, long n, int border) {
  if (i >= 0 && i < n) return i;
  switch (border) {
    case kBorderReflect:
      return i < 0 ? -i - 1 : 2 * n - i - 1;
    case kBorderWrap:
      return i < 0 ? n + i : i - n;
    case kBorderReverse:
      return i < 0 ? std::min(-i, n - 1) : 2 * n - i - 2;
    default:  // replicate
      return std::clamp(i, 0L, n - 1);
  }
}

bool is_filter_border(int border) {
  return border == kBorderReplicate || border == kBorderReflect ||
         border == kBorderWrap || border == kBorderReverse;
}

}  // namespace

extern "C" long orc_border_index(long index, long limit, int border) {
  return border_index(index, limit, border);
}

// ---------------------------------------------------------------------------
// a4: sigma half-kernel (include/kleidicv/filters/sigma.h:27-79).
// The default sigma `ksize * 0.15F + 0.35F` and `value = hk * multiplier -
// error` are single C expressions in the reference and contract to fused
// multiply-adds under clang's default fp-contract on AArch64 (the default
// sigma differs from the unfused one for 35 odd sizes; the taps change only
// for 235x235).  Parity: pinned against a double-precision model of the same
// generator (every entry equal, or at a rounding boundary) and the reference
// KATs within their stated +-1 (test_gaussian_blur.cpp:777-785).
extern "C" int orc_gaussian_half_kernel(uint8_t *half_kernel, size_t half_size,
                                        float sigma) {
  const size_t mid = half_size - 1;
  const size_t ksize = mid * 2 + 1;
  if (sigma == 0.0f) sigma = std::fmaf(static_cast<float>(ksize), 0.15f, 0.35f);
  float tmp[255];
  float coefficient = 1 / -(2 * sigma * sigma);
  float sum = 0.0f;
  size_t j = mid;
  for (size_t i = 0; i < mid; i++, j--) {
    tmp[i] = std::exp(static_cast<float>(j) * static_cast<float>(j) *
                      coefficient);
    sum += tmp[i];
  }
  float multiplier = 256 / std::fmaf(sum, 2.0f, 1.0f);
  float error = 0.0f;
  for (size_t i = 0; i < mid; i++) {
    float value = std::fmaf(tmp[i], multiplier, -error);
    float rounded = std::round(value);
    half_kernel[i] = static_cast<uint8_t>(rounded);
    error = rounded - value;
  }
  uint16_t mid_value = static_cast<uint16_t>(std::round(multiplier - error));
  if (mid_value == 256) return 0;
  half_kernel[mid] = static_cast<uint8_t>(mid_value);
  return 1;
}

// ---------------------------------------------------------------------------
// a2/a3/a4: Gaussian blur.
// Binomial taps: src/filters/gaussian_blur_fixed_neon.cpp:27-157 (3, 5),
// :160-295 (7), :297-459 (9).  The reference's two passes never overflow
// their u16/u32 accumulators, so the result is the single-rounded 2-D sum
// (sum w_i w_j src + 2^(s-1)) >> s.
// sigma != 0: :461-529 -- vertical pass rounds to u8 ((acc + 128) >> 8),
// horizontal pass repeats on the u8 intermediate; :565-573 is the "copy" case.
// Validation: src/filters/gaussian_blur_api.cpp:21-53 and
// include/kleidicv/filters/gaussian_blur.h:38-95.
// Parity: binomial pinned by ValidImageSize{3,5,7,9} KATs (exact) and by a
// brute-force 2-D convolution over padded input in tests/.
namespace {

const int kBinomial3[] = {1, 2, 1};
const int kBinomial5[] = {1, 4, 6, 4, 1};
const int kBinomial7[] = {2, 7, 14, 18, 14, 7, 2};
const int kBinomial9[] = {4, 13, 30, 51, 60, 51, 30, 13, 4};

int gaussian_validate(const uint8_t *src, const uint8_t *dst, size_t width,
                      size_t height, size_t channels, size_t k, int border) {
  if (!is_filter_border(border)) return kErrNotImplemented;
  if (k < 3 || k > 255 || (k & 1) == 0) return kErrNotImplemented;
  if (width < k - 1 || height < k - 1) return kErrNotImplemented;
  if (channels > 4) return kErrNotImplemented;
  // sizes without a fixed path: REPLICATE only, and wide enough for the
  // vector border blocks (include/kleidicv/filters/gaussian_blur.h:66-77)
  if (!(k <= 9 || k == 15 || k == 21)) {
    if (border != 1) return kErrNotImplemented;
    const size_t margin = k / 2;
    if (width < (margin + 7) / 8 * 8 + margin) return kErrNotImplemented;
  }
  if (!src || !dst) return kErrNull;
  return kErrOk;
}

}  // namespace

extern "C" int orc_gaussian_blur_u8_stripe(const uint8_t *src,
                                           size_t src_stride, uint8_t *dst,
                                           size_t dst_stride, size_t width,
                                           size_t height, size_t channels,
                                           size_t k, float sigma, int border,
                                           size_t y_begin, size_t y_end) {
  int err = gaussian_validate(src, dst, width, height, channels, k, border);
  if (err) return err;
  const long r = static_cast<long>(k / 2);
  const long W = static_cast<long>(width), H = static_cast<long>(height);
  const long C = static_cast<long>(channels);
  if (sigma == 0.0f && k <= 9) {
    const int *w = k == 3 ? kBinomial3
                          : k == 5 ? kBinomial5 : k == 7 ? kBinomial7 : kBinomial9;
    const int shift = k == 3 ? 4 : k == 5 ? 8 : k == 7 ? 12 : 16;
    for (long y = static_cast<long>(y_begin); y < static_cast<long>(y_end); ++y)
      for (long x = 0; x < W; ++x)
        for (long c = 0; c < C; ++c) {
          uint64_t acc = 0;
          for (long i = -r; i <= r; ++i) {
            const uint8_t *row = src + border_index(y + i, H, border) * src_stride;
            for (long j = -r; j <= r; ++j)
              acc += static_cast<uint64_t>(w[i + r]) * w[j + r] *
                     row[border_index(x + j, W, border) * C + c];
          }
          dst[y * dst_stride + x * C + c] =
              static_cast<uint8_t>((acc + (1ull << (shift - 1))) >> shift);
        }
    return kErrOk;
  }
  // sigma path (also 15x15 / 21x21 which have no binomial variant)
  uint8_t half[128];
  const size_t half_size = k / 2 + 1;
  if (!orc_gaussian_half_kernel(half, half_size, sigma)) {
    for (size_t y = y_begin; y < y_end; ++y)
      std::memcpy(dst + y * dst_stride, src + y * src_stride, width * channels);
    return kErrOk;
  }
  int w[255];
  for (long i = 0; i <= r; ++i) w[i] = w[2 * r - i] = half[i];
  std::vector<uint8_t> mid(width * channels);
  for (long y = static_cast<long>(y_begin); y < static_cast<long>(y_end); ++y) {
    for (long e = 0; e < W * C; ++e) {
      uint32_t acc = 0;
      for (long i = -r; i <= r; ++i)
        acc += w[i + r] * src[border_index(y + i, H, border) * src_stride + e];
      mid[e] = static_cast<uint8_t>((acc + 128) >> 8);
    }
    for (long x = 0; x < W; ++x)
      for (long c = 0; c < C; ++c) {
        uint32_t acc = 0;
        for (long j = -r; j <= r; ++j)
          acc += w[j + r] * mid[border_index(x + j, W, border) * C + c];
        dst[y * dst_stride + x * C + c] = static_cast<uint8_t>((acc + 128) >> 8);
      }
  }
  return kErrOk;
}

extern "C" int orc_gaussian_blur_u8(const uint8_t *src, size_t src_stride,
                                    uint8_t *dst, size_t dst_stride,
                                    size_t width, size_t height,
                                    size_t channels, size_t k, float sigma,
                                    int border) {
  return orc_gaussian_blur_u8_stripe(src, src_stride, dst, dst_stride, width,
                                     height, channels, k, sigma, border, 0,
                                     height);
}

// ---------------------------------------------------------------------------
// Blur and downsample (cv::pyrDown): the 5x5 binomial Gaussian evaluated at
// even source rows and columns only.  src/filters/blur_and_downsample_neon.cpp
// sums [1,4,6,4,1] vertically into u16 (max 4080) and horizontally into u16
// (max 65280), then rounds once: (acc + 128) >> 8.  Output is
// ((w+1)/2) x ((h+1)/2).  Validation order: blur_and_downsample.h:28-32
// (w, h >= 4, 1 <= channels <= 4), then the fixed border types
// (blur_and_downsample_api.cpp:25-28), then pointers and image size
// (blur_and_downsample_neon.cpp:358-360).
// Parity: pinned by the reference's Minimal_u8 KAT
// (test_blur_and_downsample.cpp:179-206) and by a numpy padded brute force
// in tests/, and equal to the binomial Gaussian 5x5 at even positions.
extern "C" int orc_blur_and_downsample_u8_stripe(
    const uint8_t *src, size_t src_stride, size_t width, size_t height,
    uint8_t *dst, size_t dst_stride, size_t channels, int border,
    size_t y_begin, size_t y_end) {
  if (width < 4 || height < 4 || channels < 1 || channels > 4)
    return kErrNotImplemented;
  if (!is_filter_border(border)) return kErrNotImplemented;
  if (!src || !dst) return kErrNull;
  static const int kTap[5] = {1, 4, 6, 4, 1};
  const long W = static_cast<long>(width), H = static_cast<long>(height);
  const long C = static_cast<long>(channels);
  const long dw = (W + 1) / 2;
  // y_begin / y_end are source rows, as the reference's stripes are
  for (long y = (static_cast<long>(y_begin) + 1) & ~1L; y < static_cast<long>(y_end); y += 2)
    for (long x = 0; x < dw; ++x)
      for (long c = 0; c < C; ++c) {
        uint32_t acc = 0;
        for (int j = 0; j < 5; ++j) {
          uint32_t col = 0;
          const long xs = border_index(2 * x + j - 2, W, border);
          for (int i = 0; i < 5; ++i)
            col += kTap[i] * src[border_index(y + i - 2, H, border) * src_stride + xs * C + c];
          acc += kTap[j] * col;
        }
        dst[(y / 2) * dst_stride + x * C + c] = static_cast<uint8_t>((acc + 128) >> 8);
      }
  return kErrOk;
}

extern "C" int orc_blur_and_downsample_u8(const uint8_t *src, size_t src_stride,
                                          size_t width, size_t height, uint8_t *dst,
                                          size_t dst_stride, size_t channels, int border) {
  return orc_blur_and_downsample_u8_stripe(src, src_stride, width, height, dst, dst_stride,
                                           channels, border, 0, height);
}

// ---------------------------------------------------------------------------
// a5: separable filter.  src/filters/separable_filter_2d_neon.cpp:19-119:
// vertical u16 saturating sum, horizontal u32 sum, saturating narrow to u8.
// With unsigned taps an inner saturation (>= 65535) forces the output to 255
// whenever its horizontal tap is non-zero, so the result equals
// min(255, sum_i sum_j ky_i kx_j src_ij).  Odd sizes other than 5 are an
// extension (the reference returns NOT_IMPLEMENTED, separable_filter_2d.h:48-60).
// Parity: pinned by the overflow KATs (test_separable_filter_2d.cpp:129-262)
// and by a brute-force sum with explicit u16 saturation in tests/.
extern "C" int orc_separable_filter_2d_u8_stripe(
    const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
    size_t width, size_t height, size_t channels, const uint8_t *kx,
    size_t kw, const uint8_t *ky, size_t kh, int border, size_t y_begin,
    size_t y_end) {
  if (kw != kh || (kw & 1) == 0 || kw < 3) return kErrNotImplemented;
  if (width < kw - 1 || height < kw - 1) return kErrNotImplemented;
  if (!is_filter_border(border)) return kErrNotImplemented;
  const long r = static_cast<long>(kw / 2);
  const long W = static_cast<long>(width), H = static_cast<long>(height);
  const long C = static_cast<long>(channels);
  for (long y = static_cast<long>(y_begin); y < static_cast<long>(y_end); ++y)
    for (long x = 0; x < W; ++x)
      for (long c = 0; c < C; ++c) {
        uint64_t acc = 0;
        for (long i = -r; i <= r; ++i) {
          const uint8_t *row = src + border_index(y + i, H, border) * src_stride;
          for (long j = -r; j <= r; ++j)
            acc += static_cast<uint64_t>(ky[i + r]) * kx[j + r] *
                   row[border_index(x + j, W, border) * C + c];
        }
        dst[y * dst_stride + x * C + c] =
            static_cast<uint8_t>(std::min<uint64_t>(acc, 255));
      }
  return kErrOk;
}

extern "C" int orc_separable_filter_2d_u8(const uint8_t *src, size_t src_stride,
                                          uint8_t *dst, size_t dst_stride,
                                          size_t width, size_t height,
                                          size_t channels, const uint8_t *kx,
                                          size_t kw, const uint8_t *ky,
                                          size_t kh, int border) {
  return orc_separable_filter_2d_u8_stripe(src, src_stride, dst, dst_stride,
                                           width, height, channels, kx, kw, ky,
                                           kh, border, 0, height);
}

// ---------------------------------------------------------------------------
// a6: median blur.  The reference's sorting networks
// (src/filters/median_blur_sorting_network_{3x3,5x5,7x7}.h) select the exact
// median; any exact selection gives the same value, so the oracle sorts.
// Borders use the fixed offsets of workspace/border_*.h through
// process_filter_2d.h:18-68.  Validation: include/kleidicv/filters/
// median_blur.h:173-195.
// Parity: pinned by brute force (the method of test_median_blur.cpp:160-190).
extern "C" int orc_median_blur_u8_stripe(const uint8_t *src, size_t src_stride,
                                         uint8_t *dst, size_t dst_stride,
                                         size_t width, size_t height,
                                         size_t channels, size_t k, int border,
                                         size_t y_begin, size_t y_end) {
  if (!src || !dst) return kErrNull;
  if (src == dst || channels > 4 || height < k - 1 || width < k - 1 ||
      k < 3 || k > 255 || (k & 1) == 0 || !is_filter_border(border))
    return kErrNotImplemented;
  const long r = static_cast<long>(k / 2);
  const long W = static_cast<long>(width), H = static_cast<long>(height);
  const long C = static_cast<long>(channels);
  std::vector<uint8_t> win(k * k);
  for (long y = static_cast<long>(y_begin); y < static_cast<long>(y_end); ++y)
    for (long x = 0; x < W; ++x)
      for (long c = 0; c < C; ++c) {
        size_t n = 0;
        for (long i = -r; i <= r; ++i) {
          const uint8_t *row = src + border_index(y + i, H, border) * src_stride;
          for (long j = -r; j <= r; ++j)
            win[n++] = row[border_index(x + j, W, border) * C + c];
        }
        std::nth_element(win.begin(), win.begin() + n / 2, win.begin() + n);
        dst[y * dst_stride + x * C + c] = win[n / 2];
      }
  return kErrOk;
}

extern "C" int orc_median_blur_u8(const uint8_t *src, size_t src_stride,
                                  uint8_t *dst, size_t dst_stride,
                                  size_t width, size_t height, size_t channels,
                                  size_t k, int border) {
  return orc_median_blur_u8_stripe(src, src_stride, dst, dst_stride, width,
                                   height, channels, k, border, 0, height);
}

// ---------------------------------------------------------------------------
// a7-a9: linear resize.
namespace {

constexpr int64_t kFixpBits = 16;
constexpr int64_t kFixpHalf = int64_t{1} << (kFixpBits - 1);

// resize_linear_generic_u8_neon.h:73-76
int64_t rounding_div(uint64_t nom, uint64_t denom) {
  return static_cast<int64_t>((nom + denom / 2) / denom);
}

// resize_linear_generic_u8_neon.h:85-88
int64_t aligned_scale(uint64_t x, uint64_t nom, uint64_t denom) {
  return rounding_div(((x << kFixpBits) + kFixpHalf) * nom, denom) - kFixpHalf +
         (1 << (kFixpBits - 9));
}

// Per-element result of one 16- or 8-lane table: the vector's load base, the
// two tbl indices and the 8-bit fraction, plus the number of bytes the
// vector load covers (tbl indices at or beyond it read 0).
struct Lane {
  int64_t base;
  uint8_t idx0, idx1;
  uint8_t frac;
  int load;
  bool direct_right;  // kSetRightmostLanes override of lane 15
};

struct GenericPlan {
  int ratio;
  bool upsize;
  bool set_rightmost;
};

// src/resize/resize_linear_api.cpp:82-216: path selection by channels and
// horizontal ratio for the generic kernels.
GenericPlan plan_generic(size_t sw, size_t dw, size_t C) {
  GenericPlan p{};
  if (C == 1) {
    if (dw * 14 > sw * 15) p = {1, true, false};
    else if (dw >= sw) p = {2, true, false};
    else if (dw * 2 >= sw) p = {2, false, false};
    else p = {3, false, false};
  } else if (C == 2) {
    if (dw * 6 > sw * 7) p = {1, true, false};
    else if (dw >= sw) p = {2, true, false};
    else if (dw * 2 >= sw) p = {2, false, false};
    else p = {3, false, false};
  } else if (C == 3) {
    if (dw * 2 > sw * 3) p = {1, true, false};
    else if (dw >= sw) p = {2, true, false};
    else {
      double inv = static_cast<double>(sw) / static_cast<double>(dw);
      if (inv < 1.8) p = {2, false, false};
      else p = {3, false, inv >= 2.8};  // resize_linear_generic_neon.cpp:39-52
    }
  } else {
    if (dw * 2 > sw * 3) p = {1, true, false};
    else if (dw >= sw) p = {2, true, false};
    else if (dw * 2 >= sw) p = {2, false, false};
    else p = {3, false, false};
  }
  return p;
}

int8_t sat_s8(int64_t v) {
  return static_cast<int8_t>(std::clamp<int64_t>(v, -128, 127));
}
int8_t sat_s8_u(uint64_t v) {
  return static_cast<int8_t>(std::min<uint64_t>(v, 127));
}

// Builds the per-destination-element column table by replaying the
// reference's table generator vector by vector and applying the row kernel's
// store order (full vectors first, then half vectors, later stores win).
// Generator: resize_linear_generic_u8_neon.h:136-401 (1, 2, 4 channels) and
// :404-703 (3 channels); store order: :730-770.
std::vector<Lane> generic_columns(size_t sw, size_t dw, size_t C,
                                  const GenericPlan &plan) {
  const int R = plan.ratio;
  const bool up = plan.upsize;
  const int kStep = 16, kHalfStep = 8;
  const size_t full_read = static_cast<size_t>(16 * R);
  std::vector<Lane> out(dw * C);

  auto scale_x = [&](uint64_t dx) -> uint32_t {
    return static_cast<uint32_t>(rounding_div((dx * sw) << kFixpBits, dw));
  };
  auto to_src_x = [&](uint64_t dx) { return aligned_scale(dx, sw, dw); };

  // calculate_num_of_vector_paths (:151-160)
  size_t two_x = (sw * C >= full_read) ? (dw * C) / (2 * kStep) : 0;
  size_t remaining = dw * C - two_x * 2 * kStep;
  size_t half_n = ((remaining + kHalfStep - 1) / kHalfStep);

  // tbl-style byte extraction from an int32 lane
  auto byte2 = [](int32_t v) { return static_cast<uint8_t>(static_cast<uint32_t>(v) >> 16); };
  auto byte1 = [](int32_t v) { return static_cast<uint8_t>(static_cast<uint32_t>(v) >> 8); };

  if (C != 3) {
    const bool load_two = (R == 3) && !up;  // kLoadTwo, C != 3
    const size_t half_read = load_two ? 32 : 16;
    int32_t vsx0[16];
    for (int l = 0; l < 16; ++l) vsx0[l] = static_cast<int32_t>(scale_x(l / C));

    auto fill_full = [&](size_t vec, int64_t sx_base, int64_t sx_fixp) {
      int32_t xfrac0 = static_cast<int32_t>(sx_fixp - (sx_base << kFixpBits));
      for (int l = 0; l < 16; ++l) {
        int32_t delta = static_cast<int32_t>(static_cast<uint32_t>(vsx0[l]) +
                                             static_cast<uint32_t>(xfrac0));
        int8_t i0 = static_cast<int8_t>(byte2(delta));
        int8_t i1 = static_cast<int8_t>(i0 + 1);
        if (up) {
          int8_t mx = static_cast<int8_t>((full_read - 1) / C);
          i0 = std::min<int8_t>(std::max<int8_t>(i0, 0), mx);
          i1 = std::min<int8_t>(std::max<int8_t>(i1, 0), mx);
        }
        uint8_t u0 = static_cast<uint8_t>(i0), u1 = static_cast<uint8_t>(i1);
        if (C > 1) {
          int sh = static_cast<int>(C / 2);
          uint8_t off = static_cast<uint8_t>(C == 4 ? (l % 4) : (l % 2));
          u0 = static_cast<uint8_t>((u0 << sh) + off);
          u1 = static_cast<uint8_t>((u1 << sh) + off);
        }
        out[vec * 16 + l] = Lane{sx_base * static_cast<int64_t>(C), u0, u1,
                                 byte1(delta), static_cast<int>(full_read),
                                 false};
      }
    };

    size_t dx = 0;
    int64_t sx_fixp = 0;
    const int64_t max_sx_full =
        static_cast<int64_t>((sw * C - std::min(sw * C, full_read)) / C);
    const int64_t step = rounding_div((sw * kStep / C) << kFixpBits, dw);
    for (size_t i = 0; i < two_x; ++i) {
      if ((i & 255) == 0) sx_fixp = to_src_x(dx);
      for (int h = 0; h < 2; ++h) {
        int64_t cand = sx_fixp >> kFixpBits;
        int64_t base = std::clamp<int64_t>(cand, 0, max_sx_full);
        fill_full(i * 2 + h, base, sx_fixp);
        sx_fixp += step;
        dx += kStep / C;
      }
    }
    sx_fixp = to_src_x(dx);
    const uint64_t one_px = static_cast<uint64_t>(rounding_div(sw << kFixpBits, dw));
    const int64_t max_sx_half = static_cast<int64_t>((sw * C - half_read) / C);
    const uint64_t max_dx_half = dw - (kHalfStep / C);
    const int64_t half_step =
        rounding_div((sw * kHalfStep / C) << kFixpBits, dw);
    for (size_t i = 0; i < half_n; ++i) {
      size_t dxp = std::min<size_t>(dx, max_dx_half);
      sx_fixp -= static_cast<int64_t>(dx - dxp) * static_cast<int64_t>(one_px);
      dx = dxp;
      int64_t cand = sx_fixp >> kFixpBits;
      int64_t base = std::clamp<int64_t>(cand, 0, max_sx_half);
      int32_t xfrac0 = static_cast<int32_t>(sx_fixp - base * (1 << kFixpBits));
      for (int l = 0; l < 8; ++l) {
        int32_t delta = static_cast<int32_t>(static_cast<uint32_t>(vsx0[l]) +
                                             static_cast<uint32_t>(xfrac0));
        int8_t i0 = static_cast<int8_t>(byte2(delta));
        int8_t i1 = static_cast<int8_t>(i0 + 1);
        if (up) {
          int8_t mx = static_cast<int8_t>((half_read - 1) / C);
          i0 = std::min<int8_t>(std::max<int8_t>(i0, 0), mx);
          i1 = std::min<int8_t>(std::max<int8_t>(i1, 0), mx);
        }
        uint8_t u0 = static_cast<uint8_t>(i0), u1 = static_cast<uint8_t>(i1);
        if (C > 1) {
          int sh = static_cast<int>(C / 2);
          uint8_t off = static_cast<uint8_t>(C == 4 ? (l % 4) : (l % 2));
          u0 = static_cast<uint8_t>((u0 << sh) + off);
          u1 = static_cast<uint8_t>((u1 << sh) + off);
        }
        out[dx * C + l] = Lane{base * static_cast<int64_t>(C), u0, u1,
                               byte1(delta), static_cast<int>(half_read), false};
      }
      dx += kHalfStep / C;
      sx_fixp += half_step;
    }
    return out;
  }

  // ---- 3 channels (:404-703) ----
  int32_t vsx[3][16];
  int8_t choff[3][16];
  {
    const int px_r[16] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5};
    const int px_g[16] = {5, 5, 6, 6, 6, 7, 7, 7, 8, 8, 8, 9, 9, 9, 10, 10};
    const int px_b[16] = {10, 11, 11, 11, 12, 12, 12, 13, 13, 13, 14, 14, 14, 15, 15, 15};
    const int co_r[16] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0};
    const int co_g[16] = {1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1};
    const int co_b[16] = {2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2};
    for (int l = 0; l < 16; ++l) {
      vsx[0][l] = static_cast<int32_t>(scale_x(px_r[l]));
      vsx[1][l] = static_cast<int32_t>(scale_x(px_g[l]));
      vsx[2][l] = static_cast<int32_t>(scale_x(px_b[l]));
      choff[0][l] = static_cast<int8_t>(co_r[l]);
      choff[1][l] = static_cast<int8_t>(co_g[l]);
      choff[2][l] = static_cast<int8_t>(co_b[l]);
    }
  }
  const int64_t one_px = rounding_div(sw << kFixpBits, dw);
  size_t dei = 0;
  int64_t sx_fixp = 0;
  const size_t nfull = two_x * 2;
  if (nfull > 0) {
    const int64_t max_base = static_cast<int64_t>(
        sw * 3 > full_read ? sw * 3 - full_read : 0);
    const int64_t triplet_step = rounding_div(sw << (kFixpBits + 4), dw);
    sx_fixp = to_src_x(0);
    unsigned cnt = 0;
    size_t handled = 0;
    auto fill = [&](size_t vec, int p) {
      int64_t sx_base = sx_fixp >> kFixpBits;
      int64_t src_base = sx_base * 3;
      int32_t xfrac0 = static_cast<int32_t>(sx_fixp & ((1 << kFixpBits) - 1));
      int8_t i0[16], i1[16];
      uint8_t fr[16];
      for (int l = 0; l < 16; ++l) {
        int32_t delta = static_cast<int32_t>(static_cast<uint32_t>(vsx[p][l]) +
                                             static_cast<uint32_t>(xfrac0));
        i0[l] = static_cast<int8_t>(byte2(delta));
        i1[l] = static_cast<int8_t>(i0[l] + 1);
        fr[l] = byte1(delta);
      }
      if (up) {
        int8_t mn = sat_s8(-sx_base);
        int8_t mx = sat_s8_u(static_cast<uint64_t>(sw) - 1 -
                             static_cast<uint64_t>(sx_base));
        for (int l = 0; l < 16; ++l) {
          i0[l] = std::min(std::max(i0[l], mn), mx);
          i1[l] = std::min(std::max(i1[l], mn), mx);
        }
      }
      for (int l = 0; l < 16; ++l) {
        i0[l] = static_cast<int8_t>(static_cast<int8_t>(i0[l] * 3) + choff[p][l]);
        i1[l] = static_cast<int8_t>(static_cast<int8_t>(i1[l] * 3) + choff[p][l]);
      }
      int8_t mn_idx = std::min(std::min(i0[0], i0[1]), i0[2]);
      int64_t final_base = std::clamp<int64_t>(src_base + mn_idx, 0, max_base);
      int8_t adj = static_cast<int8_t>(final_base - src_base);
      for (int l = 0; l < 16; ++l) {
        uint8_t u0 = static_cast<uint8_t>(static_cast<uint8_t>(i0[l]) -
                                          static_cast<uint8_t>(adj));
        uint8_t u1 = static_cast<uint8_t>(static_cast<uint8_t>(i1[l]) -
                                          static_cast<uint8_t>(adj));
        out[vec * 16 + l] = Lane{final_base, u0, u1, fr[l],
                                 static_cast<int>(full_read),
                                 plan.set_rightmost && l == 15};
      }
    };
    while (handled < nfull) {
      if (cnt == 512) {
        sx_fixp = to_src_x(dei / 3);
        cnt = 0;
      } else {
        cnt++;
      }
      fill(handled, 0);
      handled++;
      dei += kStep;
      if (handled == nfull) break;
      fill(handled, 1);
      handled++;
      dei += kStep;
      if (handled == nfull) break;
      fill(handled, 2);
      handled++;
      dei += kStep;
      sx_fixp += triplet_step;
    }
  }
  {
    const size_t half_read = up ? 16 : 32;
    const int64_t max_base =
        static_cast<int64_t>(sw * 3 > half_read ? sw * 3 - half_read : 0);
    const uint64_t max_dst_half = dw * 3 - kHalfStep;
    for (size_t i = 0; i < half_n; ++i) {
      dei = std::min<size_t>(dei, max_dst_half);
      size_t dx = dei / 3;
      unsigned inpx = static_cast<unsigned>(dei % 3);
      int64_t sxp0 = to_src_x(dx);
      int64_t sxp1 = to_src_x(dx + 1);
      int64_t sei0 = ((sxp0 >> kFixpBits) * 3) + inpx;
      int64_t sei1 = (sxp1 >> kFixpBits) * 3;
      int64_t base = std::clamp<int64_t>(std::min(sei0, sei1), 0, max_base);
      auto indices = [&](int64_t sx) {
        int64_t s0 = sx >> kFixpBits, s1 = s0 + 1;
        if (up) {
          s0 = std::clamp<int64_t>(s0, 0, static_cast<int64_t>(sw) - 1);
          s1 = std::clamp<int64_t>(s1, 0, static_cast<int64_t>(sw) - 1);
        }
        return std::pair<int64_t, int64_t>{s0 * 3, s1 * 3};
      };
      int64_t sx = sxp0;
      int64_t sei = sei0;
      unsigned j = 0;
      auto [s0, s1] = indices(sx);
      uint8_t xf = (sei < base) ? 0 : static_cast<uint8_t>((sx & 0xFFFF) >> 8);
      auto put = [&](unsigned lane, int64_t a, int64_t b, uint8_t f) {
        out[dei + lane] = Lane{base, static_cast<uint8_t>(a - base),
                               static_cast<uint8_t>(b - base), f,
                               static_cast<int>(half_read), false};
      };
      for (; j < 3 - inpx; ++j) put(j, s0 + inpx + j, s1 + inpx + j, xf);
      sx += one_px;
      sei = (sx >> kFixpBits) * 3;
      std::tie(s0, s1) = indices(sx);
      xf = static_cast<uint8_t>((sx & 0xFFFF) >> 8);
      while (j < 8) {
        for (unsigned k = 0; j < 8 && k < 3; ++j, ++k) put(j, s0 + k, s1 + k, xf);
        sx += one_px;
        sei = (sx >> kFixpBits) * 3;
        std::tie(s0, s1) = indices(sx);
        xf = (sei < base) ? 0 : static_cast<uint8_t>((sx & 0xFFFF) >> 8);
      }
      dei += kHalfStep;
    }
  }
  return out;
}

// lerp as vraddhn_u16(a << 8, (b - a) * w): exact in 16 bits for w <= 255.
inline uint8_t lerp8(int a, int b, int w) {
  return static_cast<uint8_t>((256 * a + (b - a) * w + 128) >> 8);
}

}  // namespace

extern "C" size_t orc_resize_generic_columns(size_t sw, size_t dw, size_t C,
                                             int ratio, int upsize,
                                             int32_t *left, int32_t *right,
                                             uint8_t *frac, uint8_t *zero) {
  GenericPlan plan = plan_generic(sw, dw, C);
  if (ratio > 0) plan = GenericPlan{ratio, upsize != 0, false};
  std::vector<Lane> t = generic_columns(sw, dw, C, plan);
  for (size_t e = 0; e < t.size(); ++e) {
    const Lane &L = t[e];
    left[e] = static_cast<int32_t>(L.base + L.idx0);
    right[e] = static_cast<int32_t>(L.base + L.idx1);
    frac[e] = L.frac;
    zero[e] = static_cast<uint8_t>((L.idx0 < L.load ? 0 : 1) |
                                   ((L.idx1 < L.load || L.direct_right) ? 0 : 2));
  }
  return t.size();
}

// a7 validation: include/kleidicv/resize/resize_linear.h:16-65.
static bool resize_implemented(size_t sw, size_t sh, size_t dw, size_t dh,
                               size_t C) {
  const size_t kMax = (1ull << 24) - 1;
  if (sw > kMax || sh > kMax || dw > kMax || dh > kMax || C == 0) return false;
  if (sw == 0 || sh == 0) return true;
  if (sw == 2 * dw && sh == 2 * dh && C <= 4) return true;
  if (C == 1 && ((sw * 2 == dw && sh * 2 == dh) || (sw * 4 == dw && sh * 4 == dh)))
    return true;
  if (C <= 4 && dw >= sw) return sw * C >= 16;
  if (C <= 4 && dw * 3 >= sw && dw < sw) {
    if (dw * C < 8) return false;
    if (dw * 2 >= sw && C != 3) return sw * C >= 16;
    return sw * C >= 32;
  }
  return false;
}

// a7/a8/a9 and the exact 1ch 2x/4x upsizers.
// Dispatcher: src/resize/resize_linear_api.cpp:82-247.
// Quarter: src/resize/resize_to_quarter_neon.cpp:179-213, (a+b+c+d+2)>>2.
// 2x/4x: src/resize/resize_linear_neon.cpp:61-193 and :195-...; both equal the
//   separable closed form with clamped taps (3,1)/4 resp. (7,1),(5,3),(3,5),
//   (1,7)/8 per axis, rounding once: the edge rows' 1-D formulas are that
//   closed form with the other axis' weight 1.
// Generic: the replayed column table plus the row kernel of
//   resize_linear_generic_u8_neon.h:730-870 (vertical lerp rounded to u8,
//   then horizontal lerp).
// Parity: quarter/2x/4x pinned by the ResizeLinearU8Params tables
// (test_resize_linear.cpp:1016-1187, tolerance 0) and closed forms; the
// generic table pinned against the direct-coordinate formula within the
// reference unit test's xfrac tolerance (test_resize_linear_generic_u8.cpp:
// 19-133, 339, 367-388) and bit-exact for exact ratios.
extern "C" int orc_resize_linear_u8_stripe(const uint8_t *src, size_t src_stride,
                                           size_t sw, size_t sh, uint8_t *dst,
                                           size_t dst_stride, size_t dw, size_t dh,
                                           size_t C, size_t y_begin, size_t y_end) {
  if (!resize_implemented(sw, sh, dw, dh, C)) return kErrNotImplemented;
  if (!src || !dst) return kErrNull;
  if (sw == 0 || sh == 0) return kErrOk;
  y_end = std::min(y_end, dh);
  if (sw == 2 * dw && sh == 2 * dh) {
    for (size_t y = y_begin; y < y_end; ++y)
      for (size_t x = 0; x < dw; ++x)
        for (size_t c = 0; c < C; ++c) {
          const uint8_t *r0 = src + 2 * y * src_stride;
          const uint8_t *r1 = r0 + src_stride;
          unsigned s = r0[2 * x * C + c] + r0[(2 * x + 1) * C + c] +
                       r1[2 * x * C + c] + r1[(2 * x + 1) * C + c];
          dst[y * dst_stride + x * C + c] = static_cast<uint8_t>((s + 2) >> 2);
        }
    return kErrOk;
  }
  if (C == 1 && ((sw * 2 == dw && sh * 2 == dh) || (sw * 4 == dw && sh * 4 == dh))) {
    const long f = static_cast<long>(dw / sw);  // 2 or 4
    // tap (index, weight) per axis for output coordinate d
    auto taps = [&](long d, long n, long &i0, long &i1, int &w0, int &w1) {
      const int unit = f == 2 ? 4 : 8;
      long q = d - f / 2;  // distance from the first source centre, in dst px
      if (q < 0) { i0 = i1 = 0; w0 = unit; w1 = 0; return; }
      long s = q / f, p = q % f;
      i0 = s; i1 = std::min(s + 1, n - 1);
      if (f == 2) { w0 = p == 0 ? 3 : 1; w1 = 4 - w0; }
      else { w0 = 7 - 2 * static_cast<int>(p); w1 = 8 - w0; }
      if (s + 1 > n - 1) { i1 = i0; }
    };
    const int shift = f == 2 ? 4 : 6;
    for (long y = static_cast<long>(y_begin); y < static_cast<long>(y_end); ++y) {
      long y0, y1; int wy0, wy1;
      taps(y, static_cast<long>(sh), y0, y1, wy0, wy1);
      for (long x = 0; x < static_cast<long>(dw); ++x) {
        long x0, x1; int wx0, wx1;
        taps(x, static_cast<long>(sw), x0, x1, wx0, wx1);
        const uint8_t *r0 = src + y0 * src_stride, *r1 = src + y1 * src_stride;
        int acc = wy0 * (wx0 * r0[x0] + wx1 * r0[x1]) +
                  wy1 * (wx0 * r1[x0] + wx1 * r1[x1]);
        dst[y * dst_stride + x] =
            static_cast<uint8_t>((acc + (1 << (shift - 1))) >> shift);
      }
    }
    return kErrOk;
  }
  std::vector<Lane> cols = generic_columns(sw, dw, C, plan_generic(sw, dw, C));
  for (size_t dy = y_begin; dy < y_end; ++dy) {
    int64_t sy_fixp = aligned_scale(dy, sh, dh);
    int64_t sy = sy_fixp >> kFixpBits;
    const int64_t max_sy = static_cast<int64_t>(sh) - 1;
    const uint8_t *top = src + std::clamp<int64_t>(sy, 0, max_sy) * src_stride;
    const uint8_t *bot = src + std::clamp<int64_t>(sy + 1, 0, max_sy) * src_stride;
    int yfrac = static_cast<int>((sy_fixp - sy * (int64_t{1} << kFixpBits)) >> 8);
    uint8_t *drow = dst + dy * dst_stride;
    for (size_t e = 0; e < cols.size(); ++e) {
      const Lane &L = cols[e];
      int a = L.idx0 < L.load ? top[L.base + L.idx0] : 0;
      int c = L.idx0 < L.load ? bot[L.base + L.idx0] : 0;
      int b, d;
      if (L.direct_right) {
        b = top[L.base + L.idx1];
        d = bot[L.base + L.idx1];
      } else {
        b = L.idx1 < L.load ? top[L.base + L.idx1] : 0;
        d = L.idx1 < L.load ? bot[L.base + L.idx1] : 0;
      }
      int left = lerp8(a, c, yfrac), right = lerp8(b, d, yfrac);
      drow[e] = lerp8(left, right, L.frac);
    }
  }
  return kErrOk;
}

extern "C" int orc_resize_linear_u8(const uint8_t *src, size_t src_stride,
                                    size_t sw, size_t sh, uint8_t *dst,
                                    size_t dst_stride, size_t dw, size_t dh,
                                    size_t C) {
  return orc_resize_linear_u8_stripe(src, src_stride, sw, sh, dst, dst_stride, dw, dh, C, 0,
                                     dh);
}

// ---------------------------------------------------------------------------
// a10: remap s16point5.  src/transform/remap_s16point5_neon.cpp: 1ch
// replicate :16-137, 1ch constant :321-425, 4ch coordinates :621-640 and
// :860-881, interpolation :142-192, entry :1212-1262; limits
// include/kleidicv/transform/remap.h:35-53.
// Parity: pinned by the integer closed form of test_remap.cpp:640-688 in
// tests/ (written independently there).
extern "C" int orc_remap_s16point5_u8(
    const uint8_t *src, size_t src_stride, size_t sw, size_t sh, uint8_t *dst,
    size_t dst_stride, size_t dw, size_t dh, size_t C, const int16_t *mapxy,
    size_t mapxy_stride, const uint16_t *mapfrac, size_t mapfrac_stride,
    int border, const uint8_t *border_value) {
  if (!src || !dst || !mapxy || !mapfrac) return kErrNull;
  if (border == kBorderConstant && !border_value) return kErrNull;
  if (!(src_stride <= 65535 / C && dw >= 8 && sw <= 32768 && sh <= 32768 &&
        (border == kBorderReplicate || border == kBorderConstant) &&
        (C == 1 || C == 4)))
    return kErrNotImplemented;
  const int xmax = static_cast<int>(sw) - 1, ymax = static_cast<int>(sh) - 1;
  for (size_t y = 0; y < dh; ++y) {
    const int16_t *mrow = reinterpret_cast<const int16_t *>(
        reinterpret_cast<const uint8_t *>(mapxy) + y * mapxy_stride);
    const uint16_t *frow = reinterpret_cast<const uint16_t *>(
        reinterpret_cast<const uint8_t *>(mapfrac) + y * mapfrac_stride);
    for (size_t x = 0; x < dw; ++x) {
      int16_t mx = mrow[2 * x], my = mrow[2 * x + 1];
      unsigned fx = frow[x] & 31, fy = (frow[x] >> 5) & 31;
      for (size_t c = 0; c < C; ++c) {
        unsigned a, b, cc, d;
        if (border == kBorderReplicate) {
          if (mx < 0) fx = 0;
          if (my < 0) fy = 0;
          int x0 = std::clamp<int>(mx, 0, xmax), y0 = std::clamp<int>(my, 0, ymax);
          int x1 = x0 + (x0 < xmax ? 1 : 0), y1 = y0 + (y0 < ymax ? 1 : 0);
          a = src[y0 * src_stride + x0 * C + c];
          b = src[y0 * src_stride + x1 * C + c];
          cc = src[y1 * src_stride + x0 * C + c];
          d = src[y1 * src_stride + x1 * C + c];
        } else {
          uint16_t ux0 = static_cast<uint16_t>(mx), uy0 = static_cast<uint16_t>(my);
          uint16_t ux1 = static_cast<uint16_t>(ux0 + 1), uy1 = static_cast<uint16_t>(uy0 + 1);
          auto px = [&](uint16_t xx, uint16_t yy) -> unsigned {
            if (xx < sw && yy < sh) return src[yy * src_stride + xx * C + c];
            return border_value[C == 4 ? c : 0];
          };
          a = px(ux0, uy0); b = px(ux1, uy0); cc = px(ux0, uy1); d = px(ux1, uy1);
        }
        unsigned line0 = (32 - fx) * a + fx * b;
        unsigned line1 = (32 - fx) * cc + fx * d;
        unsigned r = (line0 * (32 - fy) + line1 * fy + 512) >> 10;
        dst[y * dst_stride + x * C + c] = static_cast<uint8_t>(std::min(r, 255u));
      }
    }
  }
  return kErrOk;
}

// ---------------------------------------------------------------------------
// a11/a12 shared float gathers: src/transform/transform_neon.h:189-337
// (load_quad_pixels_*, lerp_2d) and :339-491 (nearest).  vcvtm/vcvta are
// saturating float->int conversions (NaN -> 0); lerp_2d's vmlaq_f32 is fused.
namespace {

uint32_t cvtm_u32(float f) {  // FCVTMU
  if (!(f == f)) return 0;
  float r = std::floor(f);
  if (r <= 0.0f) return 0;
  if (r >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(r);
}
int32_t cvtm_s32(float f) {  // FCVTMS
  if (!(f == f)) return 0;
  float r = std::floor(f);
  if (r >= 2147483648.0f) return 0x7FFFFFFF;
  if (r < -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(r);
}
float round_away(float f) {  // ties away from zero
  return std::round(f);
}
uint32_t cvta_u32(float f) {  // FCVTAU
  if (!(f == f)) return 0;
  float r = round_away(f);
  if (r <= 0.0f) return 0;
  if (r >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(r);
}
int32_t cvta_s32(float f) {  // FCVTAS
  if (!(f == f)) return 0;
  float r = round_away(f);
  if (r >= 2147483648.0f) return 0x7FFFFFFF;
  if (r < -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(r);
}

struct Gather {
  const uint8_t *src;
  size_t stride, C;
  uint32_t xmax, ymax;
  int border;
  const uint8_t *bv;
};

// Linear interpolation of channel c at (xf, yf).
uint8_t sample_linear(const Gather &g, float xf, float yf, size_t c) {
  float a, b, cc, d;
  float xfrac = xf - std::floor(xf), yfrac = yf - std::floor(yf);
  if (g.border == kBorderReplicate) {
    uint32_t x0 = std::min(cvtm_u32(xf), g.xmax);
    uint32_t y0 = std::min(cvtm_u32(yf), g.ymax);
    uint32_t x1 = x0 + ((xf >= 0.0f && x0 < g.xmax) ? 1 : 0);
    uint32_t y1 = y0 + ((yf >= 0.0f && y0 < g.ymax) ? 1 : 0);
    auto at = [&](uint32_t x, uint32_t y) {
      return static_cast<float>(g.src[static_cast<size_t>(y) * g.stride + x * g.C + c]);
    };
    a = at(x0, y0); b = at(x1, y0); cc = at(x0, y1); d = at(x1, y1);
  } else {
    uint32_t x0 = static_cast<uint32_t>(cvtm_s32(xf));
    uint32_t y0 = static_cast<uint32_t>(cvtm_s32(yf));
    uint32_t x1 = x0 + 1, y1 = y0 + 1;
    auto at = [&](uint32_t x, uint32_t y) {
      if (x <= g.xmax && y <= g.ymax)
        return static_cast<float>(g.src[static_cast<size_t>(y) * g.stride + x * g.C + c]);
      return static_cast<float>(g.bv[c]);
    };
    a = at(x0, y0); b = at(x1, y0); cc = at(x0, y1); d = at(x1, y1);
  }
  float line0 = std::fmaf(b - a, xfrac, a);
  float line1 = std::fmaf(d - cc, xfrac, cc);
  float res = std::fmaf(line1 - line0, yfrac, line0);
  return static_cast<uint8_t>(cvta_u32(res));
}

uint8_t sample_nearest(const Gather &g, float xf, float yf, size_t c) {
  if (g.border == kBorderReplicate) {
    uint32_t x = std::min(cvta_u32(xf), g.xmax);
    uint32_t y = std::min(cvta_u32(yf), g.ymax);
    return g.src[static_cast<size_t>(y) * g.stride + x * g.C + c];
  }
  uint32_t x = static_cast<uint32_t>(cvta_s32(xf));
  uint32_t y = static_cast<uint32_t>(cvta_s32(yf));
  if (x <= g.xmax && y <= g.ymax)
    return g.src[static_cast<size_t>(y) * g.stride + x * g.C + c];
  return g.bv[c];
}

}  // namespace

// a11: src/transform/remap_f32_neon.cpp:14-220.  Channels 3 and 4 apply the
// 1-channel arithmetic to each plane with its own border value (extension;
// the reference returns NOT_IMPLEMENTED, remap.h:55-74).
// Parity: pinned in tests/ exactly by an exact-rational model of the
// reference's operation sequence (transform_neon.h:189-337; every float step
// rounded once, AArch64 conversion rules): on dyadic maps, where the result
// is fixed by the mathematics and exact .5 ties tell the rounding modes
// apart; on NaN, infinite, negative and far-edge coordinates; and on
// arbitrary maps, including searched cases where the fused and unfused
// lerps split.  Also within the reference's +-1 of a double model
// (test_remap.cpp:1183-1244).  That the Arm build contracts vmlaq_f32 as
// here stays import-only.
extern "C" int orc_remap_f32_u8(const uint8_t *src, size_t src_stride,
                                size_t sw, size_t sh, uint8_t *dst,
                                size_t dst_stride, size_t dw, size_t dh,
                                size_t C, const float *mapx, size_t mapx_stride,
                                const float *mapy, size_t mapy_stride,
                                int interpolation, int border,
                                const uint8_t *border_value) {
  if (!src || !dst || !mapx || !mapy) return kErrNull;
  if (border == kBorderConstant && !border_value) return kErrNull;
  if (!(dw >= 4 && sw < (1u << 24) && sh < (1u << 24) && sw > 0 && sh > 0 &&
        (border == kBorderReplicate || border == kBorderConstant) && C >= 1 &&
        C <= 4))
    return kErrNotImplemented;
  uint8_t bv[4] = {0, 0, 0, 0};
  if (border == kBorderConstant) std::memcpy(bv, border_value, C);
  Gather g{src, src_stride, C, static_cast<uint32_t>(sw - 1),
           static_cast<uint32_t>(sh - 1), border, bv};
  for (size_t y = 0; y < dh; ++y) {
    const float *mxr = reinterpret_cast<const float *>(
        reinterpret_cast<const uint8_t *>(mapx) + y * mapx_stride);
    const float *myr = reinterpret_cast<const float *>(
        reinterpret_cast<const uint8_t *>(mapy) + y * mapy_stride);
    for (size_t x = 0; x < dw; ++x)
      for (size_t c = 0; c < C; ++c)
        dst[y * dst_stride + x * C + c] =
            interpolation == 1 ? sample_linear(g, mxr[x], myr[x], c)
                               : sample_nearest(g, mxr[x], myr[x], c);
  }
  return kErrOk;
}

// a12: src/transform/warp_perspective_neon.cpp:35-193 and
// transform_common.h:17-43.  Row constants T1*y+T2 etc. are fused (one C
// expression under clang's default contraction), the per-pixel tx0 + x*T0 is
// vmlaq_n_f32 (fused), 1/tw is an IEEE division, xf = tx*iw.
// Channels 2..4: per-plane extension of the 1-channel semantics.
// Parity: pinned by the SimpleResize KAT (test_warp_perspective.cpp:700-742,
// exact), exactly by the exact-rational model of the operation sequence on
// dyadic affine transforms (w = 1 or 2: every step exact) and on
// perspective ones, and within +-1 of a double-precision reference
// (:239-300).
extern "C" int orc_warp_perspective_u8_stripe(const uint8_t *src, size_t src_stride,
                                              size_t sw, size_t sh, uint8_t *dst,
                                              size_t dst_stride, size_t dw, size_t dh,
                                              const float T[9], size_t C,
                                              int interpolation, int border,
                                              const uint8_t *border_value, size_t y_begin,
                                              size_t y_end) {
  if (!(dw >= 8 && (interpolation == 0 || interpolation == 1) &&
        (border == kBorderReplicate || border == kBorderConstant) && C >= 1 &&
        C <= 4))
    return kErrNotImplemented;
  if (!src || !dst || !T) return kErrNull;
  if (border == kBorderConstant && !border_value) return kErrNull;
  if (sw >= (1u << 24) || sh >= (1u << 24) || dw >= (1u << 24) ||
      dh >= (1u << 24) || src_stride >= (1ull << 32) || sw == 0 || sh == 0)
    return kErrRange;
  uint8_t bv[4] = {0, 0, 0, 0};
  if (border == kBorderConstant) std::memcpy(bv, border_value, C);
  Gather g{src, src_stride, C, static_cast<uint32_t>(sw - 1),
           static_cast<uint32_t>(sh - 1), border, bv};
  // weight_may_be_zero (transform_common.h:17-43): T6*x + T7*y + T8 parses as
  // (T6*x + T7*y) + T8; clang fuses the first product into the first add.
  auto weight = [&](float x, float y) {
    return std::fmaf(T[6], x, T[7] * y) + T[8];
  };
  float wx0 = 0.0f, wx1 = static_cast<float>(dw - 1);
  float wy0 = 0.0f, wy1 = static_cast<float>(dh - 1);
  float mn = weight(wx0, wy0), mxw = mn;
  for (float w : {weight(wx1, wy0), weight(wx0, wy1), weight(wx1, wy1)}) {
    mn = std::min(w, mn);
    mxw = std::max(w, mxw);
  }
  const bool check_zero = mn * mxw <= 0.0f;
  for (size_t y = y_begin; y < std::min(y_end, dh); ++y) {
    float fy = static_cast<float>(y);
    float tx0 = std::fmaf(T[1], fy, T[2]);
    float ty0 = std::fmaf(T[4], fy, T[5]);
    float tw0 = std::fmaf(T[7], fy, T[8]);
    for (size_t x = 0; x < dw; ++x) {
      float fx = static_cast<float>(static_cast<uint32_t>(x));
      float tx = std::fmaf(fx, T[0], tx0);
      float ty = std::fmaf(fx, T[3], ty0);
      float tw = std::fmaf(fx, T[6], tw0);
      float iw = (check_zero && tw == 0.0f) ? 0.0f : 1.0f / tw;
      float xf = tx * iw, yf = ty * iw;
      for (size_t c = 0; c < C; ++c)
        dst[y * dst_stride + x * C + c] = interpolation == 1
                                              ? sample_linear(g, xf, yf, c)
                                              : sample_nearest(g, xf, yf, c);
    }
  }
  return kErrOk;
}

extern "C" int orc_warp_perspective_u8(const uint8_t *src, size_t src_stride,
                                       size_t sw, size_t sh, uint8_t *dst,
                                       size_t dst_stride, size_t dw, size_t dh,
                                       const float T[9], size_t C,
                                       int interpolation, int border,
                                       const uint8_t *border_value) {
  return orc_warp_perspective_u8_stripe(src, src_stride, sw, sh, dst, dst_stride, dw, dh, T, C,
                                        interpolation, border, border_value, 0, dh);
}

// ---------------------------------------------------------------------------
// a13: src/conversions/yuv420sp_to_rgb_neon.cpp:100-160 and
// yuv420_to_rgb_neon.h:17-259 with constants yuv42x_coefficients.h:95-105.
// Parity: pinned exactly by the integer formula of
// test_yuv420sp_to_rgb.cpp:112-160 in tests/, with the weights read from
// yuv42x_coefficients.h:95-105 and the test by tests/golden/extract_fixtures.py,
// on random planes and on every (U, V) pair; and on the fixed inputs (:42-53).
extern "C" int orc_yuv_semiplanar_to_rgb_u8(const uint8_t *src_y,
                                            size_t src_y_stride,
                                            const uint8_t *src_uv,
                                            size_t src_uv_stride, uint8_t *dst,
                                            size_t dst_stride, size_t width,
                                            size_t height, int fmt) {
  if ((fmt & 0x0F) != 0x02 || (fmt & 0x80)) return kErrNotImplemented;
  if (!src_y || !src_uv || !dst) return kErrNull;
  const bool vu = fmt & 0x10, bgr = fmt & 0x20, alpha = fmt & 0x40;
  const size_t oc = alpha ? 4 : 3;
  for (size_t y = 0; y < height; ++y)
    for (size_t x = 0; x < width; ++x) {
      int Y = std::max(src_y[y * src_y_stride + x] - 16, 0);
      const uint8_t *uvp = src_uv + (y / 2) * src_uv_stride + (x / 2) * 2;
      int u = uvp[vu ? 1 : 0] - 128, v = uvp[vu ? 0 : 1] - 128;
      int yw = 1220542 * Y;
      int r = (yw + 1673527 * v + (1 << 19)) >> 20;
      int gg = (yw - 409993 * u - 852492 * v + (1 << 19)) >> 20;
      int b = (yw + 2116026 * u + (1 << 19)) >> 20;
      if (bgr) std::swap(r, b);
      uint8_t *o = dst + y * dst_stride + x * oc;
      o[0] = static_cast<uint8_t>(std::clamp(r, 0, 255));
      o[1] = static_cast<uint8_t>(std::clamp(gg, 0, 255));
      o[2] = static_cast<uint8_t>(std::clamp(b, 0, 255));
      if (alpha) o[3] = 0xFF;
    }
  return kErrOk;
}

// ---------------------------------------------------------------------------
// One-buffer YUV to RGB (src/conversions/yuv_to_rgb_api.cpp:22-66).
//  - 4:2:0 planar (I420 / YV12): the chroma addressing of
//    test_yuv420p_to_rgb.cpp:161-210 (and yuv420p_to_rgb_sc.h:46-110): the
//    first chroma plane at height rows, the second at height + height/4 rows
//    plus width/2 bytes when height % 4 >= 2, each chroma row advancing by
//    width/2 and stride - width/2 in turn (the second plane taking the
//    longer step first when height % 4 == 2); pixel x reads element x >> 1.
//  - 4:2:2 interleaved (YUYV / UYVY / YVYU), even width:
//    test_yuv422_to_rgb.cpp:145-190 with yuv422_to_rgb_sc.h:398-470.
//  - 4:4:4 interleaved: full-range 14-bit constants of
//    yuv444_coefficients.h, test_yuv444_to_rgb.cpp:101-125.
// 4:2:0 and 4:2:2 use the BT.601 limited-range arithmetic of NV12 above.
// Parity: 4:2:0 and 4:2:2 are pinned in tests/ by equivalence with the
// NV12 oracle (itself pinned by the reference's fixed inputs) on the same
// chroma, 4:4:4 by the reference's fixed inputs (:43-48) and by the vector
// form of yuv444_to_rgb_sc.h (pre-shift by 2, top 16 bits of 32), written
// independently in tests/.
namespace {
void put_rgb(uint8_t *o, int r, int g, int b, bool bgr, bool alpha) {
  if (bgr) std::swap(r, b);
  o[0] = static_cast<uint8_t>(std::clamp(r, 0, 255));
  o[1] = static_cast<uint8_t>(std::clamp(g, 0, 255));
  o[2] = static_cast<uint8_t>(std::clamp(b, 0, 255));
  if (alpha) o[3] = 0xFF;
}

void bt601_limited(int Y, int u, int v, int *r, int *g, int *b) {
  const int yw = 1220542 * std::max(Y - 16, 0);
  *r = (yw + 1673527 * v + (1 << 19)) >> 20;
  *g = (yw - 409993 * u - 852492 * v + (1 << 19)) >> 20;
  *b = (yw + 2116026 * u + (1 << 19)) >> 20;
}
}  // namespace

extern "C" int orc_yuv_to_rgb_u8(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                 size_t dst_stride, size_t width, size_t height, int fmt) {
  const int base = fmt & 0x0F, flags = fmt & ~0x0F;
  const bool bgr = fmt & 0x20, alpha = fmt & 0x40;
  const size_t oc = alpha ? 4 : 3;
  if (base == 0x01) {
    if (flags & ~0x60) return kErrNotImplemented;
    if (!src || !dst) return kErrNull;
    for (size_t y = 0; y < height; ++y)
      for (size_t x = 0; x < width; ++x) {
        const uint8_t *p = src + y * src_stride + 3 * x;
        const int Y = p[0], u = p[1] - 128, v = p[2] - 128;
        put_rgb(dst + y * dst_stride + x * oc, Y + ((18678 * v + (1 << 13)) >> 14),
                Y + ((-6472 * u - 9519 * v + (1 << 13)) >> 14),
                Y + ((33292 * u + (1 << 13)) >> 14), bgr, alpha);
      }
    return kErrOk;
  }
  if (base == 0x03) {
    if (flags & ~0x70) return kErrNotImplemented;
    if (!src || !dst) return kErrNull;
    const bool yv12 = fmt & 0x10;
    const uint8_t *p1 = src + src_stride * height;
    const uint8_t *p2 = src + src_stride * (height + height / 4) + (width / 2) * ((height % 4) / 2);
    const size_t steps[2] = {width / 2, src_stride - width / 2};
    size_t i1 = 0, i2 = height % 4 == 2 ? 1 : 0;
    for (size_t y = 0; y < height; ++y) {
      for (size_t x = 0; x < width; ++x) {
        int u = p1[x >> 1] - 128, v = p2[x >> 1] - 128;
        if (yv12) std::swap(u, v);
        int r, g, b;
        bt601_limited(src[y * src_stride + x], u, v, &r, &g, &b);
        put_rgb(dst + y * dst_stride + x * oc, r, g, b, bgr, alpha);
      }
      if (y % 2 == 1) {
        p1 += steps[(i1++) & 1];
        p2 += steps[(i2++) & 1];
      }
    }
    return kErrOk;
  }
  // everything else is routed to 4:2:2, which checks pointers, then the
  // width, then the format
  if (!src || !dst) return kErrNull;
  if (width < 2 || width % 2) return kErrNotImplemented;
  const bool vu = fmt & 0x10, cf = fmt & 0x80;
  if (base != 0x04 || (vu && cf)) return kErrNotImplemented;
  const int yo = cf ? 1 : 0, uo = cf ? 0 : (vu ? 3 : 1), vo = cf ? 2 : (vu ? 1 : 3);
  for (size_t y = 0; y < height; ++y)
    for (size_t x = 0; x < width; x += 2) {
      const uint8_t *p = src + y * src_stride + 2 * x;
      for (int k = 0; k < 2; ++k) {
        int r, g, b;
        bt601_limited(p[yo + 2 * k], p[uo] - 128, p[vo] - 128, &r, &g, &b);
        put_rgb(dst + y * dst_stride + (x + k) * oc, r, g, b, bgr, alpha);
      }
    }
  return kErrOk;
}

// ---------------------------------------------------------------------------
// RGB to YUV (src/conversions/rgb_to_yuv_api.cpp:20-45 routes 4:4:4 and
// 4:2:2 by base format, everything else to planar 4:2:0).  The source has 4
// channels with the alpha flag, B first with the BGR flag.
//  - 4:2:0 (planar and semi-planar): 20-bit limited-range weights
//    (yuv42x_coefficients.h:17-29); Y = (sum + 2^19 + 16 2^20) >> 20; U and V
//    from the top-left pixel of each 2x2 block, (sum + 2^19 + 128 2^20) >> 20
//    (rgb_to_yuv420_sc.h:160-230).  Planar chroma rows: U at row h/4 of the
//    chroma area plus (w+1)/2 bytes when (h/2) is odd, V at row
//    (h + height + 1)/4 plus (w+1)/2 bytes when (h + height + 1)/2 is odd
//    (test_rgb_to_yuv420p.cpp:196-236).
//  - 4:2:2: 14-bit weights (rgb_to_yuv422_sc.h:8-19); Y per pixel, U and V
//    from the pair sums, each (sum + bias) >> 14 keeping the low byte.
//  - 4:4:4: Y = (4899 R + 9617 G + 1868 B + 2^13) >> 14, then
//    U = sat((128 2^14 + 8061 (B - Y) + 2^13) >> 14),
//    V = sat((128 2^14 + 14369 (R - Y) + 2^13) >> 14)
//    (rgb_to_yuv444_coefficients.h, rgb_to_yuv444_sc.h:18-110).
// Parity: pinned in tests/ by the reference tests' own scalar formulas for
// 4:2:0 (test_rgb_to_yuv420p.cpp:158-236, test_rgb_to_yuv420sp.cpp) and by
// round trips through the YUV to RGB oracle within the conversions' error.
namespace {
struct Px {
  int r, g, b;
};

Px px_at(const uint8_t *row, size_t x, int scn, bool bgr) {
  const uint8_t *p = row + x * scn;
  return bgr ? Px{p[2], p[1], p[0]} : Px{p[0], p[1], p[2]};
}

int y420(Px p) { return (269484 * p.r + 528482 * p.g + 102760 * p.b + (1 << 19) + (16 << 20)) >> 20; }
int u420(Px p) { return (-155188 * p.r - 305135 * p.g + 460324 * p.b + (1 << 19) + (128 << 20)) >> 20; }
int v420(Px p) { return (460324 * p.r - 385875 * p.g - 74448 * p.b + (1 << 19) + (128 << 20)) >> 20; }

bool rgb_format_420(int fmt, int base) {
  return (fmt & 0x0F) == base && (fmt & ~0x7F) == 0;
}
}  // namespace

extern "C" int orc_rgb_to_yuv_semiplanar_u8(const uint8_t *src, size_t src_stride,
                                            uint8_t *y_dst, size_t y_stride, uint8_t *uv_dst,
                                            size_t uv_stride, size_t width, size_t height,
                                            int fmt) {
  if (!src || !y_dst || !uv_dst) return kErrNull;
  if (!rgb_format_420(fmt, 0x02)) return kErrNotImplemented;
  const bool vu = fmt & 0x10, bgr = fmt & 0x20;
  const int scn = (fmt & 0x40) ? 4 : 3;
  for (size_t y = 0; y < height; ++y)
    for (size_t x = 0; x < width; ++x) {
      const Px p = px_at(src + y * src_stride, x, scn, bgr);
      y_dst[y * y_stride + x] = static_cast<uint8_t>(y420(p));
      if (y % 2 == 0 && x % 2 == 0) {
        uint8_t *uv = uv_dst + (y / 2) * uv_stride + x;
        uv[vu ? 1 : 0] = static_cast<uint8_t>(u420(p));
        uv[vu ? 0 : 1] = static_cast<uint8_t>(v420(p));
      }
    }
  return kErrOk;
}

extern "C" int orc_rgb_to_yuv_u8(const uint8_t *src, size_t src_stride, uint8_t *dst,
                                 size_t dst_stride, size_t width, size_t height, int fmt) {
  if (!src || !dst) return kErrNull;
  const int base = fmt & 0x0F;
  const bool bgr = fmt & 0x20;
  const int scn = (fmt & 0x40) ? 4 : 3;
  if (base == 0x01) {
    if (fmt & ~0x61) return kErrNotImplemented;
    for (size_t y = 0; y < height; ++y)
      for (size_t x = 0; x < width; ++x) {
        const Px p = px_at(src + y * src_stride, x, scn, bgr);
        const int Y = (4899 * p.r + 9617 * p.g + 1868 * p.b + (1 << 13)) >> 14;
        const int U = ((128 << 14) + 8061 * (p.b - Y) + (1 << 13)) >> 14;
        const int V = ((128 << 14) + 14369 * (p.r - Y) + (1 << 13)) >> 14;
        uint8_t *o = dst + y * dst_stride + 3 * x;
        o[0] = static_cast<uint8_t>(std::clamp(Y, 0, 255));
        o[1] = static_cast<uint8_t>(std::clamp(U, 0, 255));
        o[2] = static_cast<uint8_t>(std::clamp(V, 0, 255));
      }
    return kErrOk;
  }
  if (base == 0x04) {
    if (width < 2 || width % 2) return kErrNotImplemented;
    const bool vu = fmt & 0x10, cf = fmt & 0x80;
    if (vu && cf) return kErrNotImplemented;
    const int yo = cf ? 1 : 0, uo = cf ? 0 : (vu ? 3 : 1), vo = cf ? 2 : (vu ? 1 : 3);
    for (size_t y = 0; y < height; ++y)
      for (size_t x = 0; x < width; x += 2) {
        const Px p0 = px_at(src + y * src_stride, x, scn, bgr);
        const Px p1 = px_at(src + y * src_stride, x + 1, scn, bgr);
        const int yb = (1 << 13) + (16 << 14), cb = (1 << 13) + (256 << 13);
        const int r = p0.r + p1.r, g = p0.g + p1.g, b = p0.b + p1.b;
        uint8_t *o = dst + y * dst_stride + 2 * x;
        o[yo] = static_cast<uint8_t>((yb + 4211 * p0.r + 8258 * p0.g + 1606 * p0.b) >> 14);
        o[yo + 2] = static_cast<uint8_t>((yb + 4211 * p1.r + 8258 * p1.g + 1606 * p1.b) >> 14);
        o[uo] = static_cast<uint8_t>((cb - 1212 * r - 2384 * g + 3596 * b) >> 14);
        o[vo] = static_cast<uint8_t>((cb + 3596 * r - 3015 * g - 582 * b) >> 14);
      }
    return kErrOk;
  }
  if (!rgb_format_420(fmt, 0x03)) return kErrNotImplemented;
  const bool yv12 = fmt & 0x10;
  const size_t cw = (width + 1) / 2;
  uint8_t *area = dst + dst_stride * height;
  for (size_t y = 0; y < height; ++y)
    for (size_t x = 0; x < width; ++x) {
      const Px p = px_at(src + y * src_stride, x, scn, bgr);
      dst[y * dst_stride + x] = static_cast<uint8_t>(y420(p));
      if (y % 2 == 0 && x % 2 == 0) {
        const size_t hv = y + height + 1;
        uint8_t *ur = area + dst_stride * (y / 4) + ((y / 2) % 2) * cw;
        uint8_t *vr = area + dst_stride * (hv / 4) + ((hv / 2) % 2) * cw;
        int u = u420(p), v = v420(p);
        if (yv12) std::swap(u, v);
        ur[x / 2] = static_cast<uint8_t>(u);
        vr[x / 2] = static_cast<uint8_t>(v);
      }
    }
  return kErrOk;
}
