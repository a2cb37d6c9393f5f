"""Independent numpy formulations used to pin the oracle.

Each function restates what the reference's own tests compute for an op, or
what the mathematics fixes, in a form that does not share code or structure
with the oracle: borders come from numpy.pad, convolutions from
sliding-window sums, medians from numpy.median, float references in double
precision.  They run only on small inputs.
"""

import numpy as np

# KleidiCV border -> numpy.pad mode (adapters/opencv/kleidicv_hal.cpp:405-435
# and test/framework/border.cpp:117-285 give the correspondences).
PAD_MODE = {"REPLICATE": "edge", "REFLECT": "symmetric", "WRAP": "wrap",
            "REVERSE": "reflect"}

BINOMIAL = {3: ([1, 2, 1], 4), 5: ([1, 4, 6, 4, 1], 8),
            7: ([2, 7, 14, 18, 14, 7, 2], 12), 9: ([4, 13, 30, 51, 60, 51, 30, 13, 4], 16)}


def _hwc(a):
    return a if a.ndim == 3 else a[:, :, None]


def padded(src, r, border):
    s = _hwc(src)
    return np.pad(s, ((r, r), (r, r), (0, 0)), mode=PAD_MODE[border])


def windows(src, k, border):
    """(h, w, c, k, k) view of every k x k window."""
    p = padded(src, k // 2, border)
    return np.lib.stride_tricks.sliding_window_view(p, (k, k), axis=(0, 1))


def binomial_blur(src, k, border):
    taps, shift = BINOMIAL[k]
    w2 = np.outer(taps, taps).astype(np.int64)
    acc = (windows(src, k, border).astype(np.int64) * w2).sum(axis=(-1, -2))
    out = ((acc + (1 << (shift - 1))) >> shift).astype(np.uint8)
    return out.reshape(src.shape)


def blur_and_downsample(src, border):
    """cv::pyrDown as the reference's test computes it
    (test_blur_and_downsample.cpp:86-103): the 5x5 binomial at even rows and
    columns, (acc + 128) / 256."""
    return binomial_blur(src, 5, border)[::2, ::2]


def separable_two_stage(src, kx, ky, border):
    """The reference's vector path: u16-saturated vertical sum, then a
    horizontal sum saturated to u8 (separable_filter_2d_neon.cpp:37-69)."""
    k = len(kx)
    r = k // 2
    p = padded(src, r, border).astype(np.int64)
    h, w = src.shape[0], src.shape[1]
    vert = np.zeros((h, p.shape[1], p.shape[2]), np.int64)
    for i in range(k):
        vert += int(ky[i]) * p[i:i + h]
    vert = np.minimum(vert, 65535)
    horiz = np.zeros((h, w, p.shape[2]), np.int64)
    for j in range(k):
        horiz += int(kx[j]) * vert[:, j:j + w]
    return np.minimum(horiz, 255).astype(np.uint8).reshape(src.shape)


def median(src, k, border):
    win = windows(src, k, border).reshape(src.shape[0], src.shape[1], -1, k * k)
    return np.median(win, axis=-1).astype(np.uint8).reshape(src.shape)


def remap_s16point5(src, mapxy, mapfrac, border, border_value):
    """Closed form of test/api/test_remap.cpp:638-688."""
    s = _hwc(src).astype(np.int64)
    h, w, c = s.shape
    x = mapxy[..., 0].astype(np.int64)
    y = mapxy[..., 1].astype(np.int64)
    fx = (mapfrac & 31).astype(np.int64)[..., None]
    fy = ((mapfrac >> 5) & 31).astype(np.int64)[..., None]

    def get(xx, yy):
        if border == "REPLICATE":
            return s[np.clip(yy, 0, h - 1), np.clip(xx, 0, w - 1)]
        inside = (xx >= 0) & (xx < w) & (yy >= 0) & (yy < h)
        v = s[np.clip(yy, 0, h - 1), np.clip(xx, 0, w - 1)]
        bv = np.array(border_value[:c], np.int64)
        return np.where(inside[..., None], v, bv)

    a, b, cc, d = get(x, y), get(x + 1, y), get(x, y + 1), get(x + 1, y + 1)
    r = ((32 - fx) * (32 - fy) * a + fx * (32 - fy) * b + (32 - fx) * fy * cc +
         fx * fy * d + 512) // 1024
    out = r.astype(np.uint8)
    return out if src.ndim == 3 else out[..., 0]


def _double_sample(src, xf, yf, border, border_value, interpolation):
    s = _hwc(src).astype(np.float64)
    h, w, c = s.shape

    def get(xx, yy):
        if border == "REPLICATE":
            return s[np.clip(yy, 0, h - 1), np.clip(xx, 0, w - 1)]
        inside = (xx >= 0) & (xx < w) & (yy >= 0) & (yy < h)
        v = s[np.clip(yy, 0, h - 1), np.clip(xx, 0, w - 1)]
        return np.where(inside[..., None], v, np.array(border_value[:c], np.float64))

    xf = np.asarray(xf, np.float64)
    yf = np.asarray(yf, np.float64)
    if interpolation == "NEAREST":
        # round half away from zero, as FCVTA does
        xi = np.clip(np.sign(xf) * np.floor(np.abs(xf) + 0.5), -2 ** 40, 2 ** 40).astype(np.int64)
        yi = np.clip(np.sign(yf) * np.floor(np.abs(yf) + 0.5), -2 ** 40, 2 ** 40).astype(np.int64)
        out = get(xi, yi)
    else:
        x0 = np.clip(np.floor(xf), -2 ** 40, 2 ** 40).astype(np.int64)
        y0 = np.clip(np.floor(yf), -2 ** 40, 2 ** 40).astype(np.int64)
        ax = (xf - np.floor(xf))[..., None]
        ay = (yf - np.floor(yf))[..., None]
        a, b = get(x0, y0), get(x0 + 1, y0)
        cc, d = get(x0, y0 + 1), get(x0 + 1, y0 + 1)
        top = a + (b - a) * ax
        bot = cc + (d - cc) * ax
        out = np.floor(top + (bot - top) * ay + 0.5)
    out = np.clip(out, 0, 255).astype(np.uint8)
    return out if src.ndim == 3 else out[..., 0]


def remap_f32_double(src, mapx, mapy, interpolation, border, border_value):
    """Double-precision counterpart of test/api/test_remap.cpp:1183-1244."""
    return _double_sample(src, mapx, mapy, border, border_value, interpolation)


def warp_double(src, t, dst_w, dst_h, interpolation, border, border_value):
    """Double-precision homography (test/api/test_warp_perspective.cpp:122-160)."""
    T = np.asarray(t, np.float64).reshape(3, 3)
    yy, xx = np.mgrid[0:dst_h, 0:dst_w].astype(np.float64)
    tw = T[2, 0] * xx + T[2, 1] * yy + T[2, 2]
    xf = (T[0, 0] * xx + T[0, 1] * yy + T[0, 2]) / tw
    yf = (T[1, 0] * xx + T[1, 1] * yy + T[1, 2]) / tw
    return _double_sample(src, xf, yf, border, border_value, interpolation)


def nv12_float(y, uv, vu=False, bgr=False, alpha=False):
    """BT.601 limited range with the coefficients the reference rounds to
    20 bits (yuv42x_coefficients.h:46-71), evaluated in double."""
    h, w = y.shape
    Y = np.maximum(y.astype(np.float64) - 16, 0)
    U = uv[np.arange(h) // 2][:, (np.arange(w) // 2) * 2 + (1 if vu else 0)].astype(np.float64) - 128
    V = uv[np.arange(h) // 2][:, (np.arange(w) // 2) * 2 + (0 if vu else 1)].astype(np.float64) - 128
    R = 1.164 * Y + 1.596 * V
    G = 1.164 * Y - 0.391 * U - 0.813 * V
    B = 1.164 * Y + 2.018 * U
    if bgr:
        R, B = B, R
    planes = [R, G, B]
    out = np.stack([np.clip(np.floor(p + 0.5), 0, 255) for p in planes], axis=-1)
    if alpha:
        out = np.concatenate([out, np.full((h, w, 1), 255.0)], axis=-1)
    return out.astype(np.uint8)


def resize_generator_direct(C, sw, dw, R, up):
    """Direct-formula column table of the reference's generator unit test
    (test/unit_neon/test_resize_linear_generic_u8.cpp:18-136), flattened to
    one (left, right, xfrac) entry per destination element in store order."""
    def rdiv(n, d):
        return (n + d // 2) // d

    def to(dx):
        return rdiv(((dx << 16) + 32768) * sw, dw) - 32768 + 128

    def sat8(v):
        return max(-128, min(127, v))

    two_x = (dw * C) // 32 if sw * C >= 16 * R else 0
    n = dw * C
    left, right, frac = [0] * n, [0] * n, [0] * n
    di = 0
    for v in range(two_x * 2):
        base = (to(di // C) >> 16) * C + di % C
        mx = sw * C - 16 * R
        base = min(max(base, 0), mx)
        lanes = []
        for i in range(16):
            sx = to(di // C)
            inpx = di % C
            s0 = min(max(sx >> 16, 0), sw - 1)
            s1 = min(max((sx >> 16) + 1, 0), sw - 1)
            lanes.append((sat8(s0 * C + inpx - base), sat8(s1 * C + inpx - base),
                          (sx & 0xFFFF) >> 8))
            di += 1
        if C == 3:
            ml = min(lanes[0][0], lanes[1][0], lanes[2][0])
            fb = min(max(base + ml, 0), mx)
            adj = fb - base
            base = fb
            lanes = [(l0 - adj, l1 - adj, f) for l0, l1, f in lanes]
        for i, (l0, l1, f) in enumerate(lanes):
            left[v * 16 + i], right[v * 16 + i], frac[v * 16 + i] = base + l0, base + l1, f
    half = (n - di + 7) // 8
    for _ in range(half):
        di = min(di, n - 8)
        base = ((to(di // C) >> 16) * C) + ((di % C) if not up else 0)
        rs = 16 * (2 if (not up and ((R == 2 and C == 3) or R == 3)) else 1)
        base = min(max(base, 0), sw * C - rs)
        for _ in range(8):
            sx = to(di // C)
            inpx = di % C
            s0 = min(max(sx >> 16, 0), sw - 1)
            s1 = min(max((sx >> 16) + 1, 0), sw - 1)
            left[di] = base + sat8(s0 * C + inpx - base)
            right[di] = base + sat8(s1 * C + inpx - base)
            frac[di] = (sx & 0xFFFF) >> 8
            di += 1
    return np.array(left), np.array(right), np.array(frac)


def i420_buffer(yplane, uplane, vplane, stride, yv12=False):
    """OpenCV's packed I420/YV12 buffer for an even height h: the luma plane,
    then the chroma area, a sequence of h chroma rows of w/2 bytes stored two
    to a source row (slot j at (j // 2) * stride + (j % 2) * w/2): the first
    plane's h/2 rows, then the second plane's."""
    h, w = yplane.shape
    assert h % 2 == 0 and w % 2 == 0
    buf = np.zeros(((3 * h + 1) // 2, stride), np.uint8)
    buf[:h, :w] = yplane
    chroma = buf[h:].reshape(-1)
    half = w // 2
    first, second = (vplane, uplane) if yv12 else (uplane, vplane)
    for j, row in enumerate(list(first) + list(second)):
        pos = (j // 2) * stride + (j % 2) * half
        chroma[pos:pos + half] = row
    return buf


def yuv444_vector_form(src3, bgr=False, alpha=False):
    """yuv444_to_rgb_sc.h's vector arithmetic: U and V pre-shifted by 2 as
    u16, multiplied into 32 bits with the 14-bit weights and the 4x deltas,
    top 16 bits kept, added to Y, saturated."""
    Y = src3[..., 0].astype(np.int64)
    U4 = src3[..., 1].astype(np.int64) << 2
    V4 = src3[..., 2].astype(np.int64) << 2
    b_delta = (-128 * 33292 + (1 << 13)) * 4
    g_delta = (-128 * -6472 - 128 * -9519 + (1 << 13)) * 4
    r_delta = (-128 * 18678 + (1 << 13)) * 4

    def top16(x):
        x = x.astype(np.int64) & 0xFFFFFFFF
        x = np.where(x >= 1 << 31, x - (1 << 32), x)
        return x >> 16
    b = Y + top16(U4 * 33292 + b_delta)
    g = Y + top16(U4 * -6472 + V4 * -9519 + g_delta)
    r = Y + top16(V4 * 18678 + r_delta)
    if bgr:
        r, b = b, r
    ch = [r, g, b] + ([np.full_like(Y, 255)] if alpha else [])
    return np.clip(np.stack(ch, axis=-1), 0, 255).astype(np.uint8)


def rgb_to_yuv420_reference_formula(rgb):
    """Per-pixel Y and per-2x2-block (top-left) U, V as the reference's tests
    compute them (test_rgb_to_yuv420p.cpp:158-195): 20-bit weights, clamp."""
    r, g, b = (rgb[..., i].astype(np.int64) for i in range(3))
    y = np.clip((269484 * r + 528482 * g + 102760 * b + (1 << 19) + (16 << 20)) >> 20, 0, 255)
    r0, g0, b0 = r[::2, ::2], g[::2, ::2], b[::2, ::2]
    u = np.clip((-155188 * r0 - 305135 * g0 + 460324 * b0 + (1 << 19) + (128 << 20)) >> 20,
                0, 255)
    v = np.clip((460324 * r0 - 385875 * g0 - 74448 * b0 + (1 << 19) + (128 << 20)) >> 20,
                0, 255)
    return y.astype(np.uint8), u.astype(np.uint8), v.astype(np.uint8)


# ---------------------------------------------------------------------------
# NV12 / NV21 -> RGB: the integer formula of the reference's NV12 test
# (test/api/test_yuv420sp_to_rgb.cpp:112-160) with the weights read from the
# reference's files by tests/golden/extract_fixtures.py.
def nv12_reference_formula(y, uv, fmt, k):
    h, w = y.shape
    vu = bool(fmt & 0x10)
    Y = np.maximum(y.astype(np.int64) - k["y_offset"], 0)
    pair = (np.arange(w) - np.arange(w) % 2)
    rows = uv[np.arange(h) // 2].astype(np.int64)
    U = rows[:, pair + (1 if vu else 0)] - k["uv_offset"]
    V = rows[:, pair + (0 if vu else 1)] - k["uv_offset"]
    half = 1 << (k["shift"] - 1)
    R = (k["y"] * Y + k["rv"] * V + half) >> k["shift"]
    G = (k["y"] * Y + k["gu"] * U + k["gv"] * V + half) >> k["shift"]
    B = (k["y"] * Y + k["bu"] * U + half) >> k["shift"]
    if fmt & 0x20:
        R, B = B, R
    ch = [R, G, B] + ([np.full_like(R, 255)] if fmt & 0x40 else [])
    return np.clip(np.stack(ch, axis=-1), 0, 255).astype(np.uint8)


# ---------------------------------------------------------------------------
# Sigma half-kernel (include/kleidicv/filters/sigma.h:27-79) in double
# precision.  The float table can differ from it only where a value sits at
# a rounding boundary; `follow` takes the float table's choice there and
# carries on, so every later entry is still checked.
def default_sigma(ksize, fused=True):
    """sigma.h:38: ksize * 0.15F + 0.35F in float, one rounding when fused
    (clang contracts the expression on AArch64)."""
    a, b, c = (float(np.float32(v)) for v in (ksize, 0.15, 0.35))
    if fused:
        return _f32(_F(a) * _F(b) + _F(c))
    return float(np.float32(np.float32(np.float32(a) * np.float32(b)) + np.float32(c)))


def gaussian_half_kernel_double(half_size, sigma, follow=None):
    mid = half_size - 1
    ksize = 2 * mid + 1
    if sigma == 0.0:
        sigma = default_sigma(ksize)
    vals = [np.exp(-(j * j) / (2.0 * sigma * sigma)) for j in range(mid, 0, -1)]
    mult = 256.0 / (2.0 * sum(vals) + 1.0)
    out, near, err = [], [], 0.0
    for i in range(mid):
        v = vals[i] * mult - err
        r = float(np.floor(v + 0.5))
        near.append(abs(v - np.floor(v) - 0.5))
        if follow is not None and follow[i] != r:
            r = float(follow[i])
        out.append(r)
        err = r - v
    m = mult - err
    near.append(abs(m - np.floor(m) - 0.5))
    return [int(v) for v in out] + [int(np.floor(m + 0.5))], near


# ---------------------------------------------------------------------------
# Float gathers (remap f32 and warp perspective).  The reference's arithmetic
# (src/transform/transform_neon.h:189-337, warp_perspective_neon.cpp:55-102)
# evaluated exactly with rationals: each float operation rounds once to the
# nearest float32 (ties to even), and the fused forms (vmla, and the row
# terms T1*y+T2 that clang contracts on AArch64) round once for the whole
# multiply-add.  `fused=False` rounds the product separately.  The integer
# conversions follow the AArch64 instruction semantics: FCVTMU/FCVTMS floor,
# FCVTAU/FCVTAS round half away from zero, NaN to 0, saturating.
from fractions import Fraction as _F
import math as _m


def _f32(q):
    """Round a rational (or +-inf) to the nearest float32, ties to even."""
    if isinstance(q, float):
        return float(np.float32(q))
    if q == 0:
        return 0.0
    s = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if _F(2) ** e > a:
        e -= 1
    e = max(e, -126)                 # subnormals share the smallest exponent
    scale = _F(2) ** (e - 23)
    n = a / scale
    fl = n.numerator // n.denominator
    rem = n - fl
    if rem > _F(1, 2) or (rem == _F(1, 2) and fl % 2 == 1):
        fl += 1
    v = fl * scale
    if v >= _F(2) ** 128:
        return s * _m.inf
    return s * float(v)


def _q(x):
    return _F(x) if _m.isfinite(x) else x


def _add(a, b):
    if not (_m.isfinite(a) and _m.isfinite(b)):
        return a + b
    return _f32(_F(a) + _F(b))


def _mul(a, b):
    if not (_m.isfinite(a) and _m.isfinite(b)):
        return a * b
    return _f32(_F(a) * _F(b))


def _fma(a, b, c, fused):
    if not fused:
        return _add(_mul(a, b), c)
    if not (_m.isfinite(a) and _m.isfinite(b) and _m.isfinite(c)):
        return a * b + c
    return _f32(_F(a) * _F(b) + _F(c))


def _div(a, b):
    if b == 0.0 or not (_m.isfinite(a) and _m.isfinite(b)):
        return float(np.float32(np.float64(a) / np.float64(b))) if b != 0.0 else (
            _m.copysign(_m.inf, a) * _m.copysign(1.0, b) if a != 0 else _m.nan)
    return _f32(_F(a) / _F(b))


def _floor(x):
    return x if not _m.isfinite(x) else float(_m.floor(x))


def _cvt(x, signed, mode):
    lo, hi = (-(1 << 31), (1 << 31) - 1) if signed else (0, (1 << 32) - 1)
    if _m.isnan(x):
        return 0
    if _m.isinf(x):
        return hi if x > 0 else lo
    if mode == "m":
        v = _m.floor(x)
    else:  # round half away from zero
        v = _m.floor(abs(x) + 0.5) * (1 if x >= 0 else -1)
        if abs(x) % 1 == 0.5:
            v = (_m.floor(abs(x)) + 1) * (1 if x >= 0 else -1)
    return int(min(max(v, lo), hi))


def _sample(s, xf, yf, interp, border, bv, fused):
    """One destination pixel (all channels) of the reference's gather."""
    h, w, c = s.shape
    xmax, ymax = w - 1, h - 1

    def px(x, y):
        return [float(v) for v in s[y, x]]

    if interp == "NEAREST":
        if border == "REPLICATE":
            x = min(_cvt(xf, False, "a"), xmax)
            y = min(_cvt(yf, False, "a"), ymax)
            return px(x, y)
        x = _cvt(xf, True, "a") & 0xFFFFFFFF
        y = _cvt(yf, True, "a") & 0xFFFFFFFF
        return px(x, y) if x <= xmax and y <= ymax else [float(v) for v in bv[:c]]
    xfrac = _add(xf, -_floor(xf)) if _m.isfinite(xf) else _m.nan
    yfrac = _add(yf, -_floor(yf)) if _m.isfinite(yf) else _m.nan
    if border == "REPLICATE":
        x0 = min(_cvt(xf, False, "m"), xmax)
        y0 = min(_cvt(yf, False, "m"), ymax)
        x1 = x0 + 1 if (xf >= 0 and x0 < xmax) else x0
        y1 = y0 + 1 if (yf >= 0 and y0 < ymax) else y0
        a, b, cc, d = px(x0, y0), px(x1, y0), px(x0, y1), px(x1, y1)
    else:
        x0 = _cvt(xf, True, "m") & 0xFFFFFFFF
        y0 = _cvt(yf, True, "m") & 0xFFFFFFFF
        x1, y1 = (x0 + 1) & 0xFFFFFFFF, (y0 + 1) & 0xFFFFFFFF
        bvals = [float(v) for v in bv[:c]]

        def get(x, y):
            return px(x, y) if x <= xmax and y <= ymax else bvals
        a, b, cc, d = get(x0, y0), get(x1, y0), get(x0, y1), get(x1, y1)
    out = []
    for k in range(c):
        line0 = _fma(_add(b[k], -a[k]), xfrac, a[k], fused)
        line1 = _fma(_add(d[k], -cc[k]), xfrac, cc[k], fused)
        r = _fma(_add(line1, -line0), yfrac, line0, fused)
        out.append(float(_cvt(r, False, "a")))
    return out


def remap_f32_exact(src, mapx, mapy, interp, border, bv, fused=True):
    s = _hwc(src)
    dh, dw = mapx.shape
    out = np.zeros((dh, dw, s.shape[2]), np.uint8)
    for y in range(dh):
        for x in range(dw):
            out[y, x] = _sample(s, float(mapx[y, x]), float(mapy[y, x]), interp, border, bv,
                                fused)
    return out if src.ndim == 3 else out[..., 0]


def warp_exact(src, t, dst_w, dst_h, interp, border, bv, fused=True):
    """warp_perspective_neon.cpp:55-102: row terms T1*y+T2 (fused as clang
    contracts them), column terms with vmla, iw = 1/tw correctly rounded (0
    when tw == 0), xf = tx * iw."""
    T = [float(np.float32(v)) for v in np.asarray(t, np.float32).reshape(9)]
    s = _hwc(src)
    out = np.zeros((dst_h, dst_w, s.shape[2]), np.uint8)
    for y in range(dst_h):
        fy = float(y)
        tx0 = _fma(T[1], fy, T[2], fused)
        ty0 = _fma(T[4], fy, T[5], fused)
        tw0 = _fma(T[7], fy, T[8], fused)
        for x in range(dst_w):
            fx = float(x)
            tx = _fma(fx, T[0], tx0, fused)
            ty = _fma(fx, T[3], ty0, fused)
            tw = _fma(fx, T[6], tw0, fused)
            iw = 0.0 if tw == 0.0 else _div(1.0, tw)
            out[y, x] = _sample(s, _mul(tx, iw), _mul(ty, iw), interp, border, bv, fused)
    return out if src.ndim == 3 else out[..., 0]
