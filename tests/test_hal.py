"""OpenCV HAL entry points without a GPU: exports, and the adapter's policy
(adapters/opencv/kleidicv_hal.cpp) decided before any device work.

A call the HAL declines returns NOT_IMPLEMENTED (1).  A call it forwards
reaches the library, whose own validation then answers; the tests forward
null pointers there, which the library reports as NULL_POINTER and the HAL
maps to UNKNOWN (-1, kleidicv_hal.cpp:78-92), so "forwarded" and "declined"
are told apart without running a kernel."""

import ctypes as C
import os

import numpy as np

import hal_lib as H
import kleidicv_hip as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_and_c_symbols_exported():
    out = os.popen("nm -D --defined-only '%s'" % K.LIB_PATH).read().split()
    for name, mangled in H.MANGLED.items():
        assert mangled in out, name
        assert ("kleidicv_hip_hal_" + name) in out, name
    text = open(os.path.join(ROOT, "include", "kleidicv_hip", "kleidicv_hal.h")).read()
    for name in H.MANGLED:
        assert "int %s(" % name in text and "int kleidicv_hip_hal_%s(" % name in text


BUF = np.zeros(1 << 16, np.uint8)
P = BUF.ctypes.data
Q = P + 8192


def test_gaussian_policy():
    g = H.fn("gaussian_blur")
    b = H.fn("gaussian_blur_binomial")
    # forwarded: null source reaches the library
    assert g(None, 64, Q, 64, 16, 16, H.CV_8U, 1, 0, 0, 0, 0, 5, 5, 0, 0, 4) == H.UNKNOWN
    assert b(None, 64, Q, 64, 16, 16, H.CV_8U, 1, 0, 0, 0, 0, 5, 4) == H.UNKNOWN
    # declined: in place, ROI margins, depth, unknown border
    assert g(P, 64, P, 64, 16, 16, H.CV_8U, 1, 0, 0, 0, 0, 5, 5, 0, 0, 4) == H.NOT_IMPLEMENTED
    for m in ((1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1)):
        assert g(P, 64, Q, 64, 16, 16, H.CV_8U, 1, *m, 5, 5, 0, 0, 4) == H.NOT_IMPLEMENTED
        assert b(P, 64, Q, 64, 16, 16, H.CV_8U, 1, *m, 5, 4) == H.NOT_IMPLEMENTED
    assert g(P, 64, Q, 64, 16, 16, H.CV_16U, 1, 0, 0, 0, 0, 5, 5, 0, 0, 4) == H.NOT_IMPLEMENTED
    assert g(P, 64, Q, 64, 16, 16, H.CV_8U, 1, 0, 0, 0, 0, 5, 5, 0, 0, 7) == H.NOT_IMPLEMENTED
    # the library's own NOT_IMPLEMENTED (constant border) maps to 1
    assert g(P, 64, Q, 64, 16, 16, H.CV_8U, 1, 0, 0, 0, 0, 5, 5, 0, 0, 0) == H.NOT_IMPLEMENTED


def test_separable_policy():
    f = H.fn("separable_filter_2d")
    k = np.ones(5, np.uint8)
    kp = k.ctypes.data
    t8 = H.cv_type(H.CV_8U, 1)
    args = dict(src=None, ss=64, st=t8, dst=Q, ds=64, dt=t8, w=16, h=16, fw=16, fh=16, ox=0,
                oy=0, kx=kp, kxl=5, ky=kp, kyl=5, kt=t8, ax=-1, ay=-1, delta=0.0, border=1)

    def call(**kw):
        a = dict(args, **kw)
        return f(*a.values())
    assert call() == H.UNKNOWN  # forwarded
    assert call(src=P, fw=17) == H.NOT_IMPLEMENTED       # submatrix
    assert call(src=P, ox=1, fw=17) == H.NOT_IMPLEMENTED
    assert call(src=Q) == H.NOT_IMPLEMENTED              # in place
    assert call(src=P, dt=H.cv_type(H.CV_8U, 3)) == H.NOT_IMPLEMENTED
    assert call(src=P, kt=H.CV_16U) == H.NOT_IMPLEMENTED
    assert call(src=P, kt=H.cv_type(H.CV_8U, 2)) == H.NOT_IMPLEMENTED
    assert call(src=P, st=H.CV_16U, dt=H.CV_16U, kt=H.CV_16U) == H.NOT_IMPLEMENTED
    assert call(src=P, border=0) == H.NOT_IMPLEMENTED    # constant
    assert call(src=P, border=5) == H.NOT_IMPLEMENTED    # transparent
    assert call(src=P, ax=1) == H.NOT_IMPLEMENTED        # off-centre anchor
    assert call(src=P, ay=3) == H.NOT_IMPLEMENTED
    assert call(src=P, delta=1.0) == H.NOT_IMPLEMENTED
    assert call(ax=2, ay=2) == H.UNKNOWN                 # centred anchor is fine


def test_resize_policy():
    f = H.fn("resize")
    t1, t4 = H.cv_type(H.CV_8U, 1), H.cv_type(H.CV_8U, 4)
    # 0.5 x 0.5 with LINEAR or AREA is forwarded
    assert f(t4, None, 256, 64, 64, Q, 128, 32, 32, 0.5, 0.5, H.INTER_AREA) == H.UNKNOWN
    assert f(t4, None, 256, 64, 64, Q, 128, 32, 32, 0.5, 0.5, H.INTER_LINEAR) == H.UNKNOWN
    assert f(t4, P, 256, 64, 64, P, 128, 32, 32, 0.5, 0.5, H.INTER_LINEAR) == H.NOT_IMPLEMENTED
    assert f(t1, P, 64, 64, 64, Q, 128, 128, 128, 2.0, 2.0, H.INTER_AREA) == H.NOT_IMPLEMENTED
    assert f(t1, P, 64, 64, 64, Q, 128, 128, 128, 2.0, 2.0, H.INTER_CUBIC) == H.NOT_IMPLEMENTED
    # scale inconsistent with the sizes
    assert f(t1, P, 64, 64, 64, Q, 128, 127, 128, 2.0, 2.0, H.INTER_LINEAR) == H.NOT_IMPLEMENTED
    assert f(t1, None, 64, 64, 64, Q, 128, 128, 128, 2.0, 2.0, H.INTER_LINEAR) == H.UNKNOWN
    assert f(t1, None, 64, 64, 64, Q, 128, 128, 128, 0.0, 0.0, H.INTER_LINEAR) == H.UNKNOWN
    assert f(H.CV_32F, P, 256, 64, 64, Q, 512, 128, 128, 2.0, 2.0, 1) == H.NOT_IMPLEMENTED


def test_median_pyrdown_policy():
    m = H.fn("medianBlur")
    assert m(None, 64, Q, 64, 16, 16, H.CV_8U, 1, 5) == H.UNKNOWN
    assert m(P, 64, P, 64, 16, 16, H.CV_8U, 1, 5) == H.NOT_IMPLEMENTED
    assert m(P, 64, Q, 64, 16, 16, H.CV_16U, 1, 5) == H.NOT_IMPLEMENTED
    d = H.fn("pyrdown")
    assert d(None, 64, 15, 9, Q, 64, 8, 5, H.CV_8U, 1, 4) == H.UNKNOWN
    assert d(P, 64, 15, 9, Q, 64, 7, 5, H.CV_8U, 1, 4) == H.NOT_IMPLEMENTED  # dst size
    assert d(P, 64, 15, 9, Q, 64, 8, 4, H.CV_8U, 1, 4) == H.NOT_IMPLEMENTED
    assert d(P, 64, 15, 9, P, 64, 8, 5, H.CV_8U, 1, 4) == H.NOT_IMPLEMENTED
    assert d(P, 64, 15, 9, Q, 64, 8, 5, H.CV_16U, 1, 4) == H.NOT_IMPLEMENTED
    assert d(P, 64, 15, 9, Q, 64, 8, 5, H.CV_8U, 1, 9) == H.NOT_IMPLEMENTED


def test_transform_policy():
    bv = np.zeros(4, np.float64)
    t = np.eye(3, dtype=np.float64)
    rs = H.fn("remap_s16")
    t1 = H.cv_type(H.CV_8U, 1)
    assert rs(t1, None, 64, 16, 16, Q, 64, 16, 16, P, 64, P, 32, H.INTER_LINEAR, 1,
              bv.ctypes.data) == H.UNKNOWN
    assert rs(t1, P, 64, 16, 16, Q, 64, 16, 16, P, 64, P, 32, H.INTER_NEAREST, 1,
              bv.ctypes.data) == H.NOT_IMPLEMENTED
    assert rs(H.CV_16U, P, 64, 16, 16, Q, 64, 16, 16, P, 64, P, 32, H.INTER_LINEAR, 1,
              bv.ctypes.data) == H.NOT_IMPLEMENTED
    assert rs(t1, P, 64, 16, 16, Q, 64, 16, 16, P, 64, P, 32, H.INTER_LINEAR, 9,
              bv.ctypes.data) == H.NOT_IMPLEMENTED
    # the library's ALIGNMENT (odd map stride) maps to NOT_IMPLEMENTED
    assert rs(t1, P, 64, 16, 16, Q, 64, 16, 16, P, 63, P, 32, H.INTER_LINEAR, 1,
              bv.ctypes.data) == H.NOT_IMPLEMENTED
    rf = H.fn("remap_f32")
    assert rf(t1, None, 64, 16, 16, Q, 64, 16, 16, P, 64, P, 64, H.INTER_LINEAR, 1,
              bv.ctypes.data) == H.UNKNOWN
    assert rf(H.cv_type(H.CV_8U, 4), None, 64, 16, 16, Q, 64, 16, 16, P, 64, P, 64,
              H.INTER_LINEAR, 1, bv.ctypes.data) == H.NOT_IMPLEMENTED
    assert rf(t1, P, 64, 16, 16, Q, 64, 16, 16, P, 64, P, 64, H.INTER_CUBIC, 1,
              bv.ctypes.data) == H.NOT_IMPLEMENTED
    wp = H.fn("warp_perspective")
    assert wp(t1, None, 64, 16, 16, Q, 64, 16, 16, t.ctypes.data, H.INTER_LINEAR, 1,
              bv.ctypes.data) == H.UNKNOWN
    assert wp(H.cv_type(H.CV_8U, 4), P, 64, 16, 16, Q, 64, 16, 16, t.ctypes.data,
              H.INTER_LINEAR, 1, bv.ctypes.data) == H.NOT_IMPLEMENTED
    assert wp(t1, P, 64, 16, 16, Q, 64, 16, 16, t.ctypes.data, H.INTER_AREA, 1,
              bv.ctypes.data) == H.NOT_IMPLEMENTED


def test_yuv_policy():
    assert H.fn("yuv_to_bgr_sp")(None, 64, Q, 192, 16, 16, 3, False, 0) == H.UNKNOWN
    assert H.fn("yuv_to_bgr_sp_ex")(None, 64, P, 64, Q, 192, 16, 16, 3, False, 0) == H.UNKNOWN
    assert H.fn("yuv_to_bgr_p")(None, 64, Q, 192, 16, 16, 3, False, 1) == H.UNKNOWN
    assert H.fn("yuv422_to_bgr")(None, 64, Q, 192, 16, 16, 3, False, 0, 1) == H.UNKNOWN
    f = H.fn("yuv_to_bgr")
    assert f(None, 64, Q, 192, 16, 16, H.CV_8U, 3, False, False) == H.UNKNOWN
    assert f(P, 64, Q, 192, 16, 16, H.CV_8U, 3, False, True) == H.NOT_IMPLEMENTED  # CbCr
    assert f(P, 64, Q, 192, 16, 16, H.CV_16U, 3, False, False) == H.NOT_IMPLEMENTED
    assert H.fn("bgr_to_yuv420_p")(None, 64, Q, 64, 16, 16, 3, False, 2) == H.UNKNOWN
    assert H.fn("bgr_to_yuv420_sp")(None, 64, Q, 64, Q, 64, 16, 16, 3, False, 1) == H.UNKNOWN
    assert H.fn("bgr_to_yuv422")(None, 64, Q, 64, 16, 16, 3, False, 0, 0) == H.UNKNOWN
    assert H.fn("bgr_to_yuv")(P, 64, Q, 64, 16, 16, H.CV_8U, 3, False, True) == H.NOT_IMPLEMENTED
    assert H.fn("bgr_to_yuv")(None, 64, Q, 64, 16, 16, H.CV_8U, 3, False, False) == H.UNKNOWN
    # 4:2:2 with an odd width: the library declines
    assert H.fn("yuv422_to_bgr")(P, 64, Q, 192, 15, 16, 3, False, 0, 0) == H.NOT_IMPLEMENTED
