"""ctypes bindings for the CPU oracle (oracle/_build/libkcv_oracle.so).

Test infrastructure: the oracle is the checker, never the product.  Each
wrapper takes and returns numpy arrays in the reference's interleaved HWC
layout.
"""

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libkcv_oracle.so")

BORDER = {"CONSTANT": 0, "REPLICATE": 1, "REFLECT": 2, "WRAP": 3,
          "REVERSE": 4, "TRANSPARENT": 5, "NONE": 6}
INTERP = {"NEAREST": 0, "LINEAR": 1}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _lib = C.CDLL(ORACLE_SO)
        sz, u8p, vp = C.c_size_t, C.POINTER(C.c_uint8), C.c_void_p
        _lib.orc_border_index.restype = C.c_long
        _lib.orc_border_index.argtypes = [C.c_long, C.c_long, C.c_int]
        _lib.orc_resize_generic_columns.restype = sz
        _lib.orc_resize_generic_columns.argtypes = [sz, sz, sz, C.c_int, C.c_int,
                                                    vp, vp, vp, vp]
        _lib.orc_gaussian_half_kernel.argtypes = [vp, sz, C.c_float]
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else C.c_void_p(0)


def _sz(*vals):
    return [C.c_size_t(int(v)) for v in vals]


def hwc(img):
    """(h, w, c) view of a 2-D or 3-D uint8 array."""
    return img if img.ndim == 3 else img[:, :, None]


def gaussian_blur(src, kernel, sigma=0.0, border="REVERSE", y_range=None):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros_like(s)
    if y_range is None:
        err = lib().orc_gaussian_blur_u8(_p(s), *_sz(w * c), _p(dst), *_sz(w * c, w, h, c, kernel),
                                         C.c_float(sigma), C.c_int(BORDER[border]))
    else:
        err = lib().orc_gaussian_blur_u8_stripe(
            _p(s), *_sz(w * c), _p(dst), *_sz(w * c, w, h, c, kernel), C.c_float(sigma),
            C.c_int(BORDER[border]), *_sz(*y_range))
    return err, dst.reshape(src.shape)


def blur_and_downsample(src, border="REVERSE", y_range=None):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros(((h + 1) // 2, (w + 1) // 2, c), np.uint8)
    y0, y1 = (0, h) if y_range is None else y_range
    err = lib().orc_blur_and_downsample_u8_stripe(
        _p(s), *_sz(w * c, w, h), _p(dst), *_sz(dst.strides[0], c), C.c_int(BORDER[border]),
        *_sz(y0, y1))
    return err, (dst if src.ndim == 3 else dst[:, :, 0])


def yuv_to_rgb(src, width, height, fmt):
    """src: 2-D uint8 buffer whose rows are the source rows (any padding)."""
    s = np.ascontiguousarray(src, dtype=np.uint8)
    oc = 4 if fmt & 0x40 else 3
    dst = np.zeros((height, width, oc), np.uint8)
    err = lib().orc_yuv_to_rgb_u8(_p(s), *_sz(s.strides[0]), _p(dst),
                                  *_sz(dst.strides[0], width, height), C.c_int(fmt))
    return err, dst


def rgb_to_yuv(src, fmt, dst_shape=None, dst_stride=None):
    """src (h, w, c); returns (err, dst) with dst rows of dst_stride bytes."""
    s = np.ascontiguousarray(src, dtype=np.uint8)
    h, w = s.shape[:2]
    base = fmt & 0x0F
    rows, row_bytes = {0x01: (h, 3 * w), 0x04: (h, 2 * w)}.get(base, ((3 * h + 1) // 2, w))
    stride = dst_stride or row_bytes
    dst = np.zeros((rows, stride), np.uint8)
    err = lib().orc_rgb_to_yuv_u8(_p(s), *_sz(s.strides[0]), _p(dst), *_sz(stride, w, h),
                                  C.c_int(fmt))
    return err, dst


def rgb_to_yuv_semiplanar(src, fmt):
    s = np.ascontiguousarray(src, dtype=np.uint8)
    h, w = s.shape[:2]
    y = np.zeros((h, w), np.uint8)
    uv = np.zeros(((h + 1) // 2, 2 * ((w + 1) // 2)), np.uint8)
    err = lib().orc_rgb_to_yuv_semiplanar_u8(_p(s), *_sz(s.strides[0]), _p(y), *_sz(w),
                                             _p(uv), *_sz(uv.shape[1], w, h), C.c_int(fmt))
    return err, y, uv


def separable_filter_2d(src, kx, ky, border="REPLICATE"):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros_like(s)
    kx = np.ascontiguousarray(kx, dtype=np.uint8)
    ky = np.ascontiguousarray(ky, dtype=np.uint8)
    err = lib().orc_separable_filter_2d_u8(_p(s), *_sz(w * c), _p(dst), *_sz(w * c, w, h, c),
                                           _p(kx), *_sz(len(kx)), _p(ky), *_sz(len(ky)),
                                           C.c_int(BORDER[border]))
    return err, dst.reshape(src.shape)


def median_blur(src, kernel, border="REPLICATE"):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros_like(s)
    err = lib().orc_median_blur_u8(_p(s), *_sz(w * c), _p(dst), *_sz(w * c, w, h, c, kernel),
                                   C.c_int(BORDER[border]))
    return err, dst.reshape(src.shape)


def resize_linear(src, dst_w, dst_h):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros((dst_h, dst_w, c), dtype=np.uint8)
    err = lib().orc_resize_linear_u8(_p(s), *_sz(w * c, w, h), _p(dst),
                                     *_sz(dst_w * c, dst_w, dst_h, c))
    return err, (dst if src.ndim == 3 else dst[:, :, 0])


def resize_columns(src_w, dst_w, channels, ratio=0, upsize=False):
    n = dst_w * channels
    left = np.zeros(n, np.int32)
    right = np.zeros(n, np.int32)
    frac = np.zeros(n, np.uint8)
    zero = np.zeros(n, np.uint8)
    lib().orc_resize_generic_columns(src_w, dst_w, channels, int(ratio), int(upsize),
                                     _p(left), _p(right), _p(frac), _p(zero))
    return left, right, frac, zero


def remap_s16point5(src, mapxy, mapfrac, border="CONSTANT", border_value=(0, 0, 0, 0)):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dh, dw = mapfrac.shape
    dst = np.zeros((dh, dw, c), dtype=np.uint8)
    mapxy = np.ascontiguousarray(mapxy, dtype=np.int16)
    mapfrac = np.ascontiguousarray(mapfrac, dtype=np.uint16)
    bv = np.array(border_value, dtype=np.uint8)
    err = lib().orc_remap_s16point5_u8(_p(s), *_sz(w * c, w, h), _p(dst), *_sz(dw * c, dw, dh, c),
                                       _p(mapxy), *_sz(dw * 4), _p(mapfrac), *_sz(dw * 2),
                                       C.c_int(BORDER[border]), _p(bv))
    return err, (dst if src.ndim == 3 else dst[:, :, 0])


def remap_f32(src, mapx, mapy, interpolation="LINEAR", border="REPLICATE",
              border_value=(0, 0, 0, 0)):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dh, dw = mapx.shape
    dst = np.zeros((dh, dw, c), dtype=np.uint8)
    mapx = np.ascontiguousarray(mapx, dtype=np.float32)
    mapy = np.ascontiguousarray(mapy, dtype=np.float32)
    bv = np.array(border_value, dtype=np.uint8)
    err = lib().orc_remap_f32_u8(_p(s), *_sz(w * c, w, h), _p(dst), *_sz(dw * c, dw, dh, c),
                                 _p(mapx), *_sz(dw * 4), _p(mapy), *_sz(dw * 4),
                                 C.c_int(INTERP[interpolation]), C.c_int(BORDER[border]), _p(bv))
    return err, (dst if src.ndim == 3 else dst[:, :, 0])


def warp_perspective(src, transform, dst_w, dst_h, interpolation="LINEAR",
                     border="REPLICATE", border_value=(0, 0, 0, 0)):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros((dst_h, dst_w, c), dtype=np.uint8)
    t = np.ascontiguousarray(transform, dtype=np.float32)
    bv = np.array(border_value, dtype=np.uint8)
    err = lib().orc_warp_perspective_u8(_p(s), *_sz(w * c, w, h), _p(dst),
                                        *_sz(dst_w * c, dst_w, dst_h), _p(t), *_sz(c),
                                        C.c_int(INTERP[interpolation]), C.c_int(BORDER[border]),
                                        _p(bv))
    return err, (dst if src.ndim == 3 else dst[:, :, 0])


def yuv_sp_to_rgb(y, uv, color_format=0x02):
    y = np.ascontiguousarray(y, dtype=np.uint8)
    uv = np.ascontiguousarray(uv, dtype=np.uint8)
    h, w = y.shape
    oc = 4 if color_format & 0x40 else 3
    dst = np.zeros((h, w, oc), dtype=np.uint8)
    err = lib().orc_yuv_semiplanar_to_rgb_u8(_p(y), *_sz(w), _p(uv), *_sz(uv.shape[1]), _p(dst),
                                             *_sz(w * oc, w, h), C.c_int(color_format))
    return err, dst


# ---------------------------------------------------------------------------
# Full-size images: the oracle's row-stripe entry points (destination rows
# [y_begin, y_end), each reading the whole source) on a thread pool, so a
# 4K or 8K comparison takes seconds.  ctypes releases the GIL during each
# call.  The stripes are disjoint, so the result equals the whole call.
def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _striped(rows, fn):
    from concurrent.futures import ThreadPoolExecutor
    n = _threads()
    bands = [(rows * i // n, rows * (i + 1) // n) for i in range(n)]
    bands = [b for b in bands if b[1] > b[0]]
    with ThreadPoolExecutor(max_workers=n) as pool:
        errs = list(pool.map(lambda b: fn(*b), bands))
    bad = [e for e in errs if e != 0]
    return bad[0] if bad else 0


def gaussian_blur_mt(src, kernel, sigma=0.0, border="REVERSE"):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros_like(s)
    err = _striped(h, lambda a, b: lib().orc_gaussian_blur_u8_stripe(
        _p(s), *_sz(w * c), _p(dst), *_sz(w * c, w, h, c, kernel), C.c_float(sigma),
        C.c_int(BORDER[border]), *_sz(a, b)))
    return err, dst.reshape(src.shape)


def median_blur_mt(src, kernel, border="REPLICATE"):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros_like(s)
    err = _striped(h, lambda a, b: lib().orc_median_blur_u8_stripe(
        _p(s), *_sz(w * c), _p(dst), *_sz(w * c, w, h, c, kernel), C.c_int(BORDER[border]),
        *_sz(a, b)))
    return err, dst.reshape(src.shape)


def separable_filter_2d_mt(src, kx, ky, border="REPLICATE"):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros_like(s)
    kx = np.ascontiguousarray(kx, dtype=np.uint8)
    ky = np.ascontiguousarray(ky, dtype=np.uint8)
    err = _striped(h, lambda a, b: lib().orc_separable_filter_2d_u8_stripe(
        _p(s), *_sz(w * c), _p(dst), *_sz(w * c, w, h, c), _p(kx), *_sz(len(kx)), _p(ky),
        *_sz(len(ky)), C.c_int(BORDER[border]), *_sz(a, b)))
    return err, dst.reshape(src.shape)


def resize_linear_mt(src, dst_w, dst_h):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros((dst_h, dst_w, c), dtype=np.uint8)
    err = _striped(dst_h, lambda a, b: lib().orc_resize_linear_u8_stripe(
        _p(s), *_sz(w * c, w, h), _p(dst), *_sz(dst_w * c, dst_w, dst_h, c, a, b)))
    return err, (dst if src.ndim == 3 else dst[:, :, 0])


def warp_perspective_mt(src, transform, dst_w, dst_h, interpolation="LINEAR",
                        border="REPLICATE", border_value=(0, 0, 0, 0)):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dst = np.zeros((dst_h, dst_w, c), dtype=np.uint8)
    t = np.ascontiguousarray(transform, dtype=np.float32)
    bv = np.array(border_value, dtype=np.uint8)
    err = _striped(dst_h, lambda a, b: lib().orc_warp_perspective_u8_stripe(
        _p(s), *_sz(w * c, w, h), _p(dst), *_sz(dst_w * c, dst_w, dst_h), _p(t), *_sz(c),
        C.c_int(INTERP[interpolation]), C.c_int(BORDER[border]), _p(bv), *_sz(a, b)))
    return err, (dst if src.ndim == 3 else dst[:, :, 0])


def remap_s16point5_mt(src, mapxy, mapfrac, border="CONSTANT", border_value=(0, 0, 0, 0)):
    # destination rows are independent: a stripe is a row-offset view
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dh, dw = mapfrac.shape
    dst = np.zeros((dh, dw, c), dtype=np.uint8)
    mapxy = np.ascontiguousarray(mapxy, dtype=np.int16)
    mapfrac = np.ascontiguousarray(mapfrac, dtype=np.uint16)
    bv = np.array(border_value, dtype=np.uint8)
    err = _striped(dh, lambda a, b: lib().orc_remap_s16point5_u8(
        _p(s), *_sz(w * c, w, h), C.c_void_p(dst.ctypes.data + a * dw * c),
        *_sz(dw * c, dw, b - a, c), C.c_void_p(mapxy.ctypes.data + a * dw * 4), *_sz(dw * 4),
        C.c_void_p(mapfrac.ctypes.data + a * dw * 2), *_sz(dw * 2), C.c_int(BORDER[border]),
        _p(bv)))
    return err, (dst if src.ndim == 3 else dst[:, :, 0])


def remap_f32_mt(src, mapx, mapy, interpolation="LINEAR", border="REPLICATE",
                 border_value=(0, 0, 0, 0)):
    s = hwc(np.ascontiguousarray(src, dtype=np.uint8))
    h, w, c = s.shape
    dh, dw = mapx.shape
    dst = np.zeros((dh, dw, c), dtype=np.uint8)
    mapx = np.ascontiguousarray(mapx, dtype=np.float32)
    mapy = np.ascontiguousarray(mapy, dtype=np.float32)
    bv = np.array(border_value, dtype=np.uint8)
    err = _striped(dh, lambda a, b: lib().orc_remap_f32_u8(
        _p(s), *_sz(w * c, w, h), C.c_void_p(dst.ctypes.data + a * dw * c),
        *_sz(dw * c, dw, b - a, c), C.c_void_p(mapx.ctypes.data + a * dw * 4), *_sz(dw * 4),
        C.c_void_p(mapy.ctypes.data + a * dw * 4), *_sz(dw * 4),
        C.c_int(INTERP[interpolation]), C.c_int(BORDER[border]), _p(bv)))
    return err, (dst if src.ndim == 3 else dst[:, :, 0])


def yuv_sp_to_rgb_mt(y, uv, color_format=0x02):
    # a stripe of chroma rows [a, b) is luma rows [2a, 2b)
    y = np.ascontiguousarray(y, dtype=np.uint8)
    uv = np.ascontiguousarray(uv, dtype=np.uint8)
    h, w = y.shape
    oc = 4 if color_format & 0x40 else 3
    dst = np.zeros((h, w, oc), dtype=np.uint8)
    pairs = (h + 1) // 2
    err = _striped(pairs, lambda a, b: lib().orc_yuv_semiplanar_to_rgb_u8(
        C.c_void_p(y.ctypes.data + 2 * a * w), *_sz(w),
        C.c_void_p(uv.ctypes.data + a * uv.shape[1]), *_sz(uv.shape[1]),
        C.c_void_p(dst.ctypes.data + 2 * a * w * oc), *_sz(w * oc, w, min(h, 2 * b) - 2 * a),
        C.c_int(color_format)))
    return err, dst
