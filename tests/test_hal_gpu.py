"""OpenCV HAL entry points on the GPU: each forwarded call gives the same
bytes as the library's drop-in entry point it maps to, with OpenCV's
argument conventions (swapBlue, uIdx, REFLECT_101, scale factors) mapped as
adapters/opencv/kleidicv_hal.cpp maps them."""

import numpy as np
import pytest

import hal_lib as H
import kleidicv_hip as K
import oracle_lib as O

pytestmark = pytest.mark.gpu


def ptr(a):
    return a.ctypes.data


def test_hal_filters(rng):
    src = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    s = src.strides[0]
    for border, kb in (("REFLECT_101", "REVERSE"), ("REPLICATE", "REPLICATE"),
                       ("WRAP", "WRAP"), ("REFLECT", "REFLECT")):
        dst = np.zeros_like(src)
        assert H.fn("gaussian_blur_binomial")(ptr(src), s, ptr(dst), s, 53, 37, H.CV_8U, 3,
                                              0, 0, 0, 0, 5, H.BORDER[border]) == H.OK
        np.testing.assert_array_equal(dst, O.gaussian_blur(src, 5, 0.0, kb)[1])
        dst = np.zeros_like(src)
        assert H.fn("gaussian_blur")(ptr(src), s, ptr(dst), s, 53, 37, H.CV_8U, 3, 0, 0, 0, 0,
                                     7, 7, 1.5, 1.5, H.BORDER[border]) == H.OK
        np.testing.assert_array_equal(dst, K.gaussian_blur(src, 7, 1.5, kb))
    kx = np.array([1, 2, 0, 3, 1], np.uint8)
    ky = np.array([0, 1, 2, 1, 0], np.uint8)
    dst = np.zeros_like(src)
    t3 = H.cv_type(H.CV_8U, 3)
    assert H.fn("separable_filter_2d")(ptr(src), s, t3, ptr(dst), s, t3, 53, 37, 53, 37, 0, 0,
                                       ptr(kx), 5, ptr(ky), 5, H.CV_8U, -1, -1, 0.0,
                                       H.BORDER["WRAP"]) == H.OK
    np.testing.assert_array_equal(dst, O.separable_filter_2d(src, kx, ky, "WRAP")[1])
    dst = np.zeros_like(src)
    assert H.fn("medianBlur")(ptr(src), s, ptr(dst), s, 53, 37, H.CV_8U, 3, 5) == H.OK
    np.testing.assert_array_equal(dst, O.median_blur(src, 5, "REPLICATE")[1])
    down = np.zeros((19, 27, 3), np.uint8)
    assert H.fn("pyrdown")(ptr(src), s, 53, 37, ptr(down), down.strides[0], 27, 19, H.CV_8U,
                           3, H.BORDER["REFLECT_101"]) == H.OK
    np.testing.assert_array_equal(down, O.blur_and_downsample(src, "REVERSE")[1])


def test_hal_resize(rng):
    src = rng.integers(0, 256, (64, 96, 4), dtype=np.uint8)
    t4 = H.cv_type(H.CV_8U, 4)
    half = np.zeros((32, 48, 4), np.uint8)
    assert H.fn("resize")(t4, ptr(src), src.strides[0], 96, 64, ptr(half), half.strides[0],
                          48, 32, 0.5, 0.5, H.INTER_AREA) == H.OK
    np.testing.assert_array_equal(half, O.resize_linear(src, 48, 32)[1])
    up = np.zeros((128, 192, 4), np.uint8)
    assert H.fn("resize")(t4, ptr(src), src.strides[0], 96, 64, ptr(up), up.strides[0],
                          192, 128, 2.0, 2.0, H.INTER_LINEAR) == H.OK
    np.testing.assert_array_equal(up, O.resize_linear(src, 192, 128)[1])


def test_hal_transforms(rng):
    src = rng.integers(0, 256, (40, 60), dtype=np.uint8)
    bv = np.array([7.4, 0, 0, 0], np.float64)  # saturate_cast: 7
    h, w = 30, 50
    xy = np.stack([rng.integers(-3, 63, (h, w)), rng.integers(-3, 43, (h, w))],
                  axis=-1).astype(np.int16)
    fr = rng.integers(0, 1024, (h, w)).astype(np.uint16)
    dst = np.zeros((h, w), np.uint8)
    assert H.fn("remap_s16")(H.CV_8U, ptr(src), 60, 60, 40, ptr(dst), w, w, h, ptr(xy),
                             xy.strides[0], ptr(fr), fr.strides[0], H.INTER_LINEAR,
                             H.BORDER["CONSTANT"], ptr(bv)) == H.OK
    np.testing.assert_array_equal(dst, K.remap_s16point5(src, xy, fr, "CONSTANT", (7, 0, 0, 0)))
    mx = rng.uniform(-2, 62, (h, w)).astype(np.float32)
    my = rng.uniform(-2, 42, (h, w)).astype(np.float32)
    dst = np.zeros((h, w), np.uint8)
    assert H.fn("remap_f32")(H.CV_8U, ptr(src), 60, 60, 40, ptr(dst), w, w, h, ptr(mx),
                             mx.strides[0], ptr(my), my.strides[0], H.INTER_LINEAR,
                             H.BORDER["REPLICATE"], ptr(bv)) == H.OK
    np.testing.assert_array_equal(dst, K.remap_f32(src, mx, my, "LINEAR", "REPLICATE"))
    t = np.array([[0.9, 0.1, 2.0], [-0.1, 0.95, 1.0], [1e-4, -1e-4, 1.0]], np.float64)
    dst = np.zeros((h, w), np.uint8)
    assert H.fn("warp_perspective")(H.CV_8U, ptr(src), 60, 60, 40, ptr(dst), w, w, h, ptr(t),
                                    H.INTER_NEAREST, H.BORDER["CONSTANT"], ptr(bv)) == H.OK
    np.testing.assert_array_equal(
        dst, K.warp_perspective(src, t.astype(np.float32), w, h, "NEAREST", "CONSTANT",
                                (7, 0, 0, 0)))


def test_hal_yuv(rng):
    h, w = 22, 34
    y = rng.integers(0, 256, (h, w), dtype=np.uint8)
    uv = rng.integers(0, 256, (h // 2, w), dtype=np.uint8)
    sp = np.concatenate([y, uv])  # one buffer: chroma after luma
    for swap_blue, dcn, u_idx in ((False, 3, 0), (True, 4, 1)):
        fmt = 0x02 | (0x20 if not swap_blue else 0) | (0x40 if dcn == 4 else 0) | \
            (0x10 if u_idx else 0)
        want = O.yuv_sp_to_rgb(y, uv, fmt)[1]
        dst = np.zeros((h, w, dcn), np.uint8)
        assert H.fn("yuv_to_bgr_sp")(ptr(sp), w, ptr(dst), dst.strides[0], w, h, dcn,
                                     swap_blue, u_idx) == H.OK
        np.testing.assert_array_equal(dst, want)
        dst = np.zeros((h, w, dcn), np.uint8)
        assert H.fn("yuv_to_bgr_sp_ex")(ptr(y), w, ptr(uv), w, ptr(dst), dst.strides[0], w, h,
                                        dcn, swap_blue, u_idx) == H.OK
        np.testing.assert_array_equal(dst, want)
    p = rng.integers(0, 256, ((3 * h + 1) // 2, w), dtype=np.uint8)
    dst = np.zeros((h, w, 3), np.uint8)
    assert H.fn("yuv_to_bgr_p")(ptr(p), w, ptr(dst), 3 * w, w, h, 3, False, 1) == H.OK
    np.testing.assert_array_equal(dst, O.yuv_to_rgb(p, w, h, 0x33)[1])  # YV12 -> BGR
    i422 = rng.integers(0, 256, (h, 2 * w), dtype=np.uint8)
    dst = np.zeros((h, w, 4), np.uint8)
    assert H.fn("yuv422_to_bgr")(ptr(i422), 2 * w, ptr(dst), 4 * w, w, h, 4, True, 0,
                                 1) == H.OK
    np.testing.assert_array_equal(dst, O.yuv_to_rgb(i422, w, h, 0xC4)[1])  # UYVY -> RGBA
    i444 = rng.integers(0, 256, (h, 3 * w), dtype=np.uint8)
    dst = np.zeros((h, w, 3), np.uint8)
    assert H.fn("yuv_to_bgr")(ptr(i444), 3 * w, ptr(dst), 3 * w, w, h, H.CV_8U, 3, False,
                              False) == H.OK
    np.testing.assert_array_equal(dst, O.yuv_to_rgb(i444, w, h, 0x21)[1])  # -> BGR


def test_hal_bgr_to_yuv(rng):
    h, w = 22, 34
    src = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    # swapBlue false: the source is BGR (kleidicv_hal.cpp:342-385)
    dst = np.zeros(((3 * h + 1) // 2, w), np.uint8)
    assert H.fn("bgr_to_yuv420_p")(ptr(src), 3 * w, ptr(dst), w, w, h, 3, False, 2) == H.OK
    np.testing.assert_array_equal(dst, O.rgb_to_yuv(src, 0x33)[1])   # BGR -> YV12
    y = np.zeros((h, w), np.uint8)
    uv = np.zeros((h // 2, w), np.uint8)
    assert H.fn("bgr_to_yuv420_sp")(ptr(src), 3 * w, ptr(y), w, ptr(uv), w, w, h, 3, True,
                                    1) == H.OK
    _, wy, wuv = O.rgb_to_yuv_semiplanar(src, 0x02)                   # RGB -> NV12
    np.testing.assert_array_equal(y, wy)
    np.testing.assert_array_equal(uv, wuv)
    dst = np.zeros((h, 3 * w), np.uint8)
    assert H.fn("bgr_to_yuv")(ptr(src), 3 * w, ptr(dst), 3 * w, w, h, H.CV_8U, 3, True,
                              False) == H.OK
    np.testing.assert_array_equal(dst, O.rgb_to_yuv(src, 0x01)[1])    # RGB -> YUV444
    dst = np.zeros((h, 2 * w), np.uint8)
    assert H.fn("bgr_to_yuv422")(ptr(src), 3 * w, ptr(dst), 2 * w, w, h, 3, False, 1,
                                 0) == H.OK
    np.testing.assert_array_equal(dst, O.rgb_to_yuv(src, 0x34)[1])    # BGR -> YVYU
