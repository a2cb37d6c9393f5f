"""C-ABI checks that need no GPU: the library loads, exports every entry
point that include/kleidicv_hip/kleidicv.h declares, and rejects invalid
arguments with the reference's error codes before touching the device."""

import ctypes as C
import os
import re

import numpy as np
import pytest

import kleidicv_hip as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kleidicv_hip", "kleidicv.h")


def declared_symbols():
    text = open(HEADER).read()
    funcs = re.findall(r"^(?:kleidicv_error_t|void \*|void|size_t)\s*\**\s*(kleidicv_\w+)\(",
                       text, re.M)
    data = re.findall(r"^extern kleidicv_error_t \(\*(kleidicv_\w+)\)", text, re.M)
    return sorted(set(funcs)), sorted(set(data))


def exported():
    out = os.popen("nm -D --defined-only '%s'" % K.LIB_PATH).read()
    syms = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 3:
            syms[parts[2]] = parts[1]
    return syms


def test_library_exports_every_declared_symbol():
    funcs, data = declared_symbols()
    assert len(funcs) >= 20 and len(data) == 4
    syms = exported()
    for f in funcs:
        assert syms.get(f) == "T", f
    for d in data:
        assert syms.get(d) in ("D", "B"), d
    lib = K.lib()
    for f in funcs:
        getattr(lib, f)
    for d in data:
        assert C.c_void_p.in_dll(lib, d).value


def test_function_pointer_symbols_are_initialised():
    lib = K.lib()
    for name in ("kleidicv_yuv_semiplanar_to_rgb_u8", "kleidicv_remap_s16point5_u8",
                 "kleidicv_remap_f32_u8"):
        assert C.c_void_p.in_dll(lib, name).value != 0


SME_TWINS = ["kleidicv_yuv_to_rgb_u8", "kleidicv_rgb_to_yuv_u8",
             "kleidicv_rgb_to_yuv_semiplanar_u8", "kleidicv_separable_filter_2d_u8",
             "kleidicv_gaussian_blur_u8", "kleidicv_blur_and_downsample_u8",
             "kleidicv_median_blur_u8", "kleidicv_resize_linear_u8"]


def test_sme_twins_exported_with_the_same_behaviour():
    # kleidicv.h:1583,1647,1744,1816,2097,2165,2220,2273,2489: the `_sme`
    # backend-selection names keep the semantics of the unsuffixed ones
    funcs, data = declared_symbols()
    syms = exported()
    for name in SME_TWINS:
        assert name + "_sme" in funcs, name
        assert syms.get(name + "_sme") == "T", name
    assert "kleidicv_yuv_semiplanar_to_rgb_u8_sme" in data
    lib = K.lib()
    a = C.c_void_p.in_dll(lib, "kleidicv_yuv_semiplanar_to_rgb_u8").value
    b = C.c_void_p.in_dll(lib, "kleidicv_yuv_semiplanar_to_rgb_u8_sme").value
    assert a == b != 0
    # same error codes for the same invalid arguments (no device needed)
    buf = np.zeros(64 * 64, np.uint8)
    p = buf.ctypes.data
    for suffix in ("", "_sme"):
        g = getattr(lib, "kleidicv_gaussian_blur_u8" + suffix)
        args = [C.c_void_p(p), C.c_size_t(64), C.c_void_p(p + 1), C.c_size_t(64),
                C.c_size_t(16), C.c_size_t(16), C.c_size_t(1), C.c_size_t(5), C.c_size_t(5),
                C.c_float(0), C.c_float(0)]
        assert g(*args, C.c_int(0)) == K.NOT_IMPLEMENTED
        assert g(C.c_void_p(0), *args[1:], C.c_int(1)) == K.NULL_POINTER
        m = getattr(lib, "kleidicv_median_blur_u8" + suffix)
        assert m(C.c_void_p(p), C.c_size_t(64), C.c_void_p(p), C.c_size_t(64), C.c_size_t(16),
                 C.c_size_t(16), C.c_size_t(1), C.c_size_t(5), C.c_size_t(5),
                 C.c_int(1)) == K.NOT_IMPLEMENTED  # in place
        r = getattr(lib, "kleidicv_resize_linear_u8" + suffix)
        assert r(C.c_void_p(0), C.c_size_t(16), C.c_size_t(16), C.c_size_t(16),
                 C.c_void_p(p), C.c_size_t(16), C.c_size_t(8), C.c_size_t(8),
                 C.c_size_t(1)) == K.NULL_POINTER


def _gauss(src, sstride, dst, dstride, w, h, c, kw, kh, sx, sy, border):
    return K.lib().kleidicv_gaussian_blur_u8(
        C.c_void_p(src), C.c_size_t(sstride), C.c_void_p(dst), C.c_size_t(dstride),
        C.c_size_t(w), C.c_size_t(h), C.c_size_t(c), C.c_size_t(kw), C.c_size_t(kh),
        C.c_float(sx), C.c_float(sy), C.c_int(border))


def test_gaussian_error_codes():
    buf = np.zeros(64 * 64, np.uint8)
    p = buf.ctypes.data
    # border first (gaussian_blur_api.cpp:28-31)
    for b in (0, 5, 6):
        assert _gauss(p, 64, p + 1, 64, 16, 16, 1, 5, 5, 0, 0, b) == K.NOT_IMPLEMENTED
    assert _gauss(p, 64, p + 1, 64, 16, 16, 1, 5, 3, 0, 0, 1) == K.NOT_IMPLEMENTED  # kw != kh
    assert _gauss(p, 64, p + 1, 64, 16, 16, 1, 4, 4, 0, 0, 1) == K.NOT_IMPLEMENTED  # even
    assert _gauss(p, 64, p + 1, 64, 3, 16, 1, 5, 5, 0, 0, 1) == K.NOT_IMPLEMENTED   # w < k-1
    assert _gauss(p, 64, p + 1, 64, 16, 16, 5, 5, 5, 0, 0, 1) == K.NOT_IMPLEMENTED  # channels
    assert _gauss(p, 64, p + 1, 64, 16, 16, 1, 5, 5, 1.0, 2.0, 1) == K.NOT_IMPLEMENTED
    assert _gauss(0, 64, p + 1, 64, 16, 16, 1, 5, 5, 0, 0, 1) == K.NULL_POINTER
    assert _gauss(p, 64, 0, 64, 16, 16, 1, 5, 5, 0, 0, 1) == K.NULL_POINTER
    assert _gauss(p, 64, p + 1, 64, 1 << 40, 1 << 40, 1, 5, 5, 0, 0, 1) == K.RANGE


def test_median_error_codes():
    lib = K.lib()
    buf = np.zeros(64 * 64, np.uint8)
    p = buf.ctypes.data

    def med(src, dst, w, h, c, k, b):
        return lib.kleidicv_median_blur_u8(C.c_void_p(src), C.c_size_t(64), C.c_void_p(dst),
                                           C.c_size_t(64), C.c_size_t(w), C.c_size_t(h),
                                           C.c_size_t(c), C.c_size_t(k), C.c_size_t(k),
                                           C.c_int(b))
    assert med(0, p, 16, 16, 1, 5, 1) == K.NULL_POINTER
    assert med(p, p, 16, 16, 1, 5, 1) == K.NOT_IMPLEMENTED  # in place
    assert med(p, p + 1, 16, 16, 1, 5, 0) == K.NOT_IMPLEMENTED  # constant border
    assert med(p, p + 1, 16, 16, 1, 4, 1) == K.NOT_IMPLEMENTED


def test_blur_and_downsample_error_codes():
    # test_blur_and_downsample.cpp:208-287, in the reference's check order
    lib = K.lib()
    buf = np.zeros(64 * 64, np.uint8)
    p = buf.ctypes.data

    def bd(src, w, h, dst, c, b):
        return lib.kleidicv_blur_and_downsample_u8(
            C.c_void_p(src), C.c_size_t(1), C.c_size_t(w), C.c_size_t(h), C.c_void_p(dst),
            C.c_size_t(1), C.c_size_t(c), C.c_int(b))
    for b in (0, 5, 6):
        assert bd(p, 4, 4, p + 1, 1, b) == K.NOT_IMPLEMENTED
    assert bd(p, 3, 3, p + 1, 1, 1) == K.NOT_IMPLEMENTED
    assert bd(p, 3, 4, p + 1, 1, 1) == K.NOT_IMPLEMENTED
    assert bd(p, 4, 3, p + 1, 1, 1) == K.NOT_IMPLEMENTED
    assert bd(p, 4, 4, p + 1, 5, 2) == K.NOT_IMPLEMENTED
    assert bd(p, 4, 4, p + 1, 0, 2) == K.NOT_IMPLEMENTED
    assert bd(0, 4, 4, p + 1, 1, 1) == K.NULL_POINTER
    assert bd(p, 4, 4, 0, 1, 1) == K.NULL_POINTER
    assert bd(p, (1 << 48) // 4 + 1, 4, p + 1, 1, 2) == K.RANGE
    assert bd(p, 1 << 48, 1 << 48, p + 1, 1, 2) == K.RANGE


def test_yuv_to_rgb_error_codes():
    # test_yuv420p_to_rgb.cpp:120-153 and the 4:2:2 / 4:4:4 equivalents
    lib = K.lib()
    buf = np.zeros(64 * 64, np.uint8)
    p = buf.ctypes.data

    def yuv(src, dst, w, h, fmt):
        return lib.kleidicv_yuv_to_rgb_u8(C.c_void_p(src), C.c_size_t(64), C.c_void_p(dst),
                                          C.c_size_t(256), C.c_size_t(w), C.c_size_t(h),
                                          C.c_int(fmt))
    for fmt in (0x03, 0x73, 0x04, 0x84, 0x14, 0x01, 0x61):
        assert yuv(0, p, 20, 10, fmt) == K.NULL_POINTER
        assert yuv(p, 0, 20, 10, fmt) == K.NULL_POINTER
        assert yuv(p, p, (1 << 48), 1, fmt) == K.RANGE
        assert yuv(p, p, 1 << 48, 1 << 48, fmt) == K.RANGE
    for fmt in (0x03, 0x01):
        assert yuv(p, p, 0, 1, fmt) == K.OK
        assert yuv(p, p, 1, 0, fmt) == K.OK
    assert yuv(p, p, 20, 10, 0x00) == K.NOT_IMPLEMENTED
    assert yuv(p, p, 20, 10, 0x83) == K.NOT_IMPLEMENTED
    assert yuv(p, p, 20, 10, 0x11) == K.NOT_IMPLEMENTED
    assert yuv(p, p, 20, 10, 0x94) == K.NOT_IMPLEMENTED
    assert yuv(p, p, 20, 10, 0x02) == K.NOT_IMPLEMENTED
    assert yuv(p, p, 19, 10, 0x04) == K.NOT_IMPLEMENTED
    assert yuv(p, p, 0, 10, 0x04) == K.NOT_IMPLEMENTED
    assert yuv(0, p, 20, 10, 0x83) == K.NOT_IMPLEMENTED  # format before pointers


def test_rgb_to_yuv_error_codes():
    # test_rgb_to_yuv420p.cpp, test_rgb_to_yuv422.cpp, test_rgb_to_yuv444.cpp
    # unsupported-argument cases
    lib = K.lib()
    buf = np.zeros(64 * 64, np.uint8)
    p = buf.ctypes.data

    def ryuv(src, dst, w, h, fmt):
        return lib.kleidicv_rgb_to_yuv_u8(C.c_void_p(src), C.c_size_t(256), C.c_void_p(dst),
                                          C.c_size_t(64), C.c_size_t(w), C.c_size_t(h),
                                          C.c_int(fmt))

    def rsp(src, y, uv, w, h, fmt):
        return lib.kleidicv_rgb_to_yuv_semiplanar_u8(
            C.c_void_p(src), C.c_size_t(256), C.c_void_p(y), C.c_size_t(64), C.c_void_p(uv),
            C.c_size_t(64), C.c_size_t(w), C.c_size_t(h), C.c_int(fmt))
    for fmt in (0x03, 0x73, 0x04, 0x84, 0x01, 0x61):
        assert ryuv(0, p, 20, 10, fmt) == K.NULL_POINTER
        assert ryuv(p, 0, 20, 10, fmt) == K.NULL_POINTER
        assert ryuv(p, p, 1 << 48, 1 << 48, fmt) == K.RANGE
    assert ryuv(p, p, 19, 10, 0x04) == K.NOT_IMPLEMENTED
    assert ryuv(p, p, 20, 10, 0x94) == K.NOT_IMPLEMENTED
    assert ryuv(p, p, 20, 10, 0x11) == K.NOT_IMPLEMENTED
    assert ryuv(p, p, 20, 10, 0x02) == K.NOT_IMPLEMENTED
    assert ryuv(p, p, 20, 10, 0x83) == K.NOT_IMPLEMENTED
    assert ryuv(0, p, 20, 10, 0x83) == K.NULL_POINTER  # pointers before the format
    assert ryuv(p, p, 0, 1, 0x03) == K.OK
    for fmt in (0x02, 0x72):
        assert rsp(0, p, p, 20, 10, fmt) == K.NULL_POINTER
        assert rsp(p, 0, p, 20, 10, fmt) == K.NULL_POINTER
        assert rsp(p, p, 0, 20, 10, fmt) == K.NULL_POINTER
    assert rsp(p, p, p, 20, 10, 0x03) == K.NOT_IMPLEMENTED
    assert rsp(p, p, p, 20, 10, 0x82) == K.NOT_IMPLEMENTED


def test_resize_error_codes():
    lib = K.lib()
    buf = np.zeros(4096, np.uint8)
    p = buf.ctypes.data

    def rs(src, sw, sh, dst, dw, dh, c):
        return lib.kleidicv_resize_linear_u8(C.c_void_p(src), C.c_size_t(sw * max(c, 1)),
                                             C.c_size_t(sw), C.c_size_t(sh), C.c_void_p(dst),
                                             C.c_size_t(dw * max(c, 1)), C.c_size_t(dw),
                                             C.c_size_t(dh), C.c_size_t(c))
    assert rs(p, 48, 2, p, 32, 1, 5) == K.NOT_IMPLEMENTED   # test_resize_linear.cpp:996-1014
    assert rs(p, 48, 2, p, 32, 1, 0) == K.NOT_IMPLEMENTED
    assert rs(p, 48, 2, p, 15, 1, 1) == K.NOT_IMPLEMENTED   # below 1/3
    assert rs(0, 32, 2, p, 16, 1, 1) == K.NULL_POINTER
    assert rs(p, 0, 0, 0, 0, 0, 1) == K.NULL_POINTER
    assert rs(p, 0, 0, p, 0, 0, 1) == K.OK


def test_remap_error_codes():
    buf = np.zeros(4096, np.uint8)
    p = buf.ctypes.data
    fn = K.data_function("kleidicv_remap_s16point5_u8", *K.REMAP_S16_ARGTYPES)
    assert fn(p, 16, 16, 4, p, 16, 16, 4, 1, p, 64, p, 32, 0, None) == K.NULL_POINTER
    assert fn(p, 16, 16, 4, p, 16, 16, 4, 1, p, 63, p, 32, 1, None) == K.ALIGNMENT
    assert fn(p, 16, 16, 4, p, 16, 16, 4, 2, p, 64, p, 32, 1, None) == K.NOT_IMPLEMENTED
    assert fn(p, 16, 16, 4, p, 7, 7, 4, 1, p, 64, p, 32, 1, None) == K.NOT_IMPLEMENTED
    # an empty source is declined before any launch (the reference reads out
    # of bounds there), for both borders and channel counts
    for c in (1, 4):
        for border in (0, 1):
            assert fn(p, 64, 0, 4, p, 64, 16, 4, c, p, 64, p, 32, border, p) == K.NOT_IMPLEMENTED
            assert fn(p, 64, 16, 0, p, 64, 16, 4, c, p, 64, p, 32, border, p) == K.NOT_IMPLEMENTED
    # with no destination rows there is nothing to read: OK, as the reference
    assert fn(p, 64, 0, 0, p, 64, 16, 0, 1, p, 64, p, 32, 1, p) == K.OK


def test_warp_error_codes():
    buf = np.zeros(4096, np.uint8)
    p = buf.ctypes.data
    t = np.eye(3, dtype=np.float32)
    lib = K.lib()

    def warp(src, sw, sh, dst, dw, dh, tp, c, interp, border, bv):
        return lib.kleidicv_warp_perspective_u8(
            C.c_void_p(src), C.c_size_t(max(sw, 1)), C.c_size_t(sw), C.c_size_t(sh),
            C.c_void_p(dst), C.c_size_t(dw), C.c_size_t(dw), C.c_size_t(dh),
            C.c_void_p(tp), C.c_size_t(c), C.c_int(interp), C.c_int(border), C.c_void_p(bv))
    tp = t.ctypes.data
    assert warp(p, 16, 16, p, 7, 4, tp, 1, 1, 1, None) == K.NOT_IMPLEMENTED  # dst_w < 8
    assert warp(p, 16, 16, p, 16, 4, tp, 1, 2, 1, None) == K.NOT_IMPLEMENTED
    assert warp(p, 16, 16, p, 16, 4, tp, 5, 1, 1, None) == K.NOT_IMPLEMENTED
    assert warp(p, 16, 16, p, 16, 4, None, 1, 1, 1, None) == K.NULL_POINTER
    assert warp(p, 16, 16, p, 16, 4, tp, 1, 1, 0, None) == K.NULL_POINTER
    assert warp(p, 0, 16, p, 16, 4, tp, 1, 1, 1, None) == K.RANGE


def test_yuv_error_codes():
    buf = np.zeros(4096, np.uint8)
    p = buf.ctypes.data
    fn = K.data_function("kleidicv_yuv_semiplanar_to_rgb_u8", *K.YUV_SP_ARGTYPES)
    assert fn(p, 16, p, 16, p, 48, 16, 4, 0x03) == K.NOT_IMPLEMENTED  # I420 is not semi-planar
    assert fn(p, 16, p, 16, p, 48, 16, 4, 0x82) == K.NOT_IMPLEMENTED
    assert fn(None, 16, p, 16, p, 48, 16, 4, 0x02) == K.NULL_POINTER
    assert fn(p, 16, None, 16, p, 48, 16, 4, 0x02) == K.NULL_POINTER


def test_thread_api_requires_scheduler():
    class MT(C.Structure):
        _fields_ = [("parallel", C.c_void_p), ("parallel_data", C.c_void_p)]
    buf = np.zeros(4096, np.uint8)
    p = buf.ctypes.data
    r = K.lib().kleidicv_thread_median_blur_u8(
        C.c_void_p(p), C.c_size_t(16), C.c_void_p(p + 1), C.c_size_t(16), C.c_size_t(16),
        C.c_size_t(16), C.c_size_t(1), C.c_size_t(5), C.c_size_t(5), C.c_int(1), MT(None, None))
    assert r == K.NULL_POINTER


LAYOUT_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "kleidicv_hip/kleidicv.h"
#define F(T, m) printf(#T " " #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("kleidicv_hip_op_desc size %zu\n", sizeof(kleidicv_hip_op_desc));
  printf("kleidicv_hip_batch_tuning size %zu\n", sizeof(kleidicv_hip_batch_tuning));
  %s
  return 0;
}
"""


def test_batch_structs_match_the_c_layout(tmp_path):
    fields = ["F(kleidicv_hip_op_desc, %s)" % n for n, _ in K.OpDesc._fields_]
    fields += ["F(kleidicv_hip_batch_tuning, %s)" % n for n, _ in K.Tuning._fields_]
    src = tmp_path / "probe.c"
    src.write_text(LAYOUT_PROBE.replace("%s", "\n  ".join(fields)))
    exe = tmp_path / "probe"
    rc = os.system("gcc -std=c99 -I'%s' '%s' -o '%s'" % (os.path.join(ROOT, "include"), src, exe))
    assert rc == 0
    want = {}
    for line in os.popen(str(exe)).read().splitlines():
        t, m, v = line.split()
        want[(t, m)] = int(v)
    assert want[("kleidicv_hip_op_desc", "size")] == C.sizeof(K.OpDesc)
    assert want[("kleidicv_hip_batch_tuning", "size")] == C.sizeof(K.Tuning)
    for n, _ in K.OpDesc._fields_:
        assert want[("kleidicv_hip_op_desc", n)] == getattr(K.OpDesc, n).offset, n
    for n, _ in K.Tuning._fields_:
        assert want[("kleidicv_hip_batch_tuning", n)] == getattr(K.Tuning, n).offset, n


def test_batch_create_argument_errors():
    lib = K.lib()
    h = C.c_void_p(0)
    assert lib.kleidicv_hip_batch_create(None, C.c_size_t(0), None, None, None) == \
        K.NULL_POINTER
    assert lib.kleidicv_hip_batch_create(None, C.c_size_t(3), None, C.byref(h), None) == \
        K.NULL_POINTER
    bad = K.Tuning(0, 0, 1, 1)  # blocks_per_cu must be 1..16
    assert lib.kleidicv_hip_batch_create(None, C.c_size_t(0), C.byref(bad), C.byref(h),
                                         None) == K.RANGE
    assert lib.kleidicv_hip_batch_run(None, None) == K.NULL_POINTER
    assert lib.kleidicv_hip_batch_tune(None, None, C.c_uint(1), None, None) == K.NULL_POINTER
    lib.kleidicv_hip_batch_destroy(None)  # no-op
