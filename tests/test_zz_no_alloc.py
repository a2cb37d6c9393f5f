"""The no-allocation resize entry point against the allocating one.

kleidicv_hip_resize_linear_u8_noalloc keeps its coordinate tables in a
fixed device array instead of a per-call allocation.  Called through the C
ABI twice in a row (the second call smaller, so stale table entries of the
first would show), it must write exactly what kleidicv_hip_resize_linear_u8
writes, without any library allocation (the allocation test hook is armed
throughout and would make an allocation fail), and it must reject tables
larger than its array.
"""

import ctypes as C

import numpy as np
import pytest

import kleidicv_hip as K

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAX_BYTES = 4 << 20


def dev():
    return torch.device("cuda:0")


def call(fn, src, dst, stream):
    h, w, c = src.shape
    dh, dw, _ = dst.shape
    return fn(C.c_void_p(src.data_ptr()), C.c_size_t(src.stride(0)), C.c_size_t(w),
              C.c_size_t(h), C.c_void_p(dst.data_ptr()), C.c_size_t(dst.stride(0)),
              C.c_size_t(dw), C.c_size_t(dh), C.c_size_t(c), C.c_void_p(stream.cuda_stream))


# (src w, h, dst w, h, channels): generic up and down paths, the exact 1-channel
# 2x path (tables) and the quarter and exact 4-channel 2x paths (no tables)
CASES = [
    [(1000, 700, 1531, 1013, 3), (400, 300, 611, 457, 3)],
    [(1920, 1080, 1280, 720, 4), (640, 360, 500, 281, 4)],
    [(700, 500, 1400, 1000, 1), (300, 200, 600, 400, 1)],
    [(3840, 2160, 1920, 1080, 4), (1000, 600, 500, 300, 4)],
    [(1024, 512, 2048, 1024, 4), (256, 128, 512, 256, 4)],
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d" % (c[0][0], c[0][1]))
def test_noalloc_resize_matches_allocating_entry_point(case):
    L = K.lib()
    stream = torch.cuda.Stream()
    g = torch.Generator(device=dev())
    g.manual_seed(81)
    for sw, sh, dw, dh, c in case:
        src = torch.randint(0, 256, (sh, sw, c), dtype=torch.uint8, device=dev(), generator=g)
        want = torch.full((dh, dw, c), 0x5A, dtype=torch.uint8, device=dev())
        got = torch.full((dh, dw, c), 0xA5, dtype=torch.uint8, device=dev())
        stream.wait_stream(torch.cuda.current_stream())
        assert call(L.kleidicv_hip_resize_linear_u8, src, want, stream) == K.OK
        torch.cuda.synchronize()
        K.fail_device_allocation(0)  # any library allocation now fails
        try:
            assert call(L.kleidicv_hip_resize_linear_u8_noalloc, src, got, stream) == K.OK
            torch.cuda.synchronize()
        finally:
            K.fail_device_allocation(-1)
        assert torch.equal(got, want), "%dx%dx%d -> %dx%d" % (sw, sh, c, dw, dh)


def test_noalloc_resize_rejects_tables_beyond_its_array():
    L = K.lib()
    stream = torch.cuda.Stream()
    # generic 4-channel upscale (about 2x) of wide rows: the tables need
    # (dst_w * 4 + dst_h) * 8 bytes, just above the array
    sw, dh = 65000, 3  # not an exact 2x (that path needs no tables)
    dw = (MAX_BYTES // 8 - dh) // 4 + 1
    src = torch.randint(0, 256, (dh, sw, 4), dtype=torch.uint8, device=dev())
    dst = torch.zeros((dh, dw, 4), dtype=torch.uint8, device=dev())
    torch.cuda.synchronize()
    assert call(L.kleidicv_hip_resize_linear_u8_noalloc, src, dst, stream) == K.RANGE
    # one destination column less fits exactly
    assert ((dw - 1) * 4 + dh) * 8 <= MAX_BYTES < (dw * 4 + dh) * 8
    ok = torch.zeros((dh, dw - 1, 4), dtype=torch.uint8, device=dev())
    want = torch.ones_like(ok)
    assert call(L.kleidicv_hip_resize_linear_u8_noalloc, src, ok, stream) == K.OK
    assert call(L.kleidicv_hip_resize_linear_u8, src, want, stream) == K.OK
    torch.cuda.synchronize()
    assert torch.equal(ok, want)
