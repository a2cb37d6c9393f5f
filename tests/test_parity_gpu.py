"""Bit-exact parity of the HIP library against the CPU oracle.

Every case runs the product through its C ABI (host buffers through the
drop-in `kleidicv_*` entry points, device buffers through `kleidicv_hip_*`)
on seeded inputs and compares with the oracle at tolerance 0.  Float ops
(remap f32, warp) are compared bit-exactly too: both sides implement the
reference's fused multiply-adds and correctly rounded division.
"""

import json
import os

import numpy as np
import pytest

import kleidicv_hip as K
import oracle_lib as O

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
BORDERS = ["REPLICATE", "REFLECT", "WRAP", "REVERSE"]


def assert_same(got, want, what=""):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    if not np.array_equal(got, want):
        bad = np.argwhere(got != want)
        i = tuple(bad[0])
        raise AssertionError("%s: %d mismatches, first at %s: got %d want %d" %
                             (what, len(bad), i, got[i], want[i]))


def padded_rows(rng, h, w, c, pad):
    """(h, w, c) view into a buffer whose rows carry `pad` extra bytes."""
    buf = rng.integers(0, 256, (h, w * c + pad), dtype=np.uint8)
    return buf[:, :w * c].reshape(h, w, c) if c > 1 else buf[:, :w]


# ---------------------------------------------------------------------------
# Gaussian
GAUSS_SIZES = [(4, 4), (9, 5), (33, 70), (257, 40), (300, 19), (513, 37)]


@pytest.mark.parametrize("k", [3, 5, 7, 9])
@pytest.mark.parametrize("border", BORDERS)
def test_gaussian_binomial(rng, k, border):
    for (w, h) in GAUSS_SIZES:
        if w < k - 1 or h < k - 1:
            continue
        for c in (1, 2, 3, 4):
            src = rng.integers(0, 256, (h, w, c), dtype=np.uint8)
            got = K.gaussian_blur(src, k, 0.0, border)
            _, want = O.gaussian_blur(src, k, 0.0, border)
            assert_same(got, want, "k%d %s %dx%dx%d" % (k, border, w, h, c))


@pytest.mark.parametrize("k", [3, 5, 7])
def test_gaussian_binomial_extreme_values(rng, k):
    # all-255 images and 0/255 patterns drive the column and row sums to
    # their largest values (255 * 16 * 16 + 128 for 5x5, the top of the
    # 16-bit horizontal pass; 32-bit sums for 7x7)
    h, w = 37, 300
    for c in (1, 2, 3, 4):
        for name, src in (("full", np.full((h, w, c), 255, np.uint8)),
                          ("stripes", np.where((np.arange(w) // 3) % 2 == 0, 255, 0)
                           .astype(np.uint8)[None, :, None].repeat(h, 0).repeat(c, 2)),
                          ("sparse", np.where(rng.random((h, w, c)) < 0.9, 255, 0).astype(np.uint8))):
            for border in ("REPLICATE", "REVERSE"):
                got = K.gaussian_blur(src, k, 0.0, border)
                _, want = O.gaussian_blur(src, k, 0.0, border)
                assert_same(got, want, "k%d %s %s c%d" % (k, name, border, c))


@pytest.mark.parametrize("k,sigma", [(3, 0.8), (5, 2.25), (7, 1.3), (9, 3.0), (15, 0.0),
                                     (21, 0.0), (15, 4.0), (21, 6.5), (5, 0.1)])
def test_gaussian_sigma(rng, k, sigma):
    for border in BORDERS:
        for (w, h) in [(k - 1, k - 1), (130, 45), (301, 23)]:
            c = int(rng.integers(1, 5))
            src = rng.integers(0, 256, (h, w, c), dtype=np.uint8)
            got = K.gaussian_blur(src, k, sigma, border)
            _, want = O.gaussian_blur(src, k, sigma, border)
            assert_same(got, want, "k%d s%.2f %s %dx%dx%d" % (k, sigma, border, w, h, c))


@pytest.mark.parametrize("k,sigma", [(11, 0.0), (13, 2.0), (17, 0.0), (25, 3.5), (31, 0.0),
                                     (63, 9.0), (127, 0.0), (255, 0.0), (255, 0.3)])
def test_gaussian_arbitrary_size(rng, k, sigma):
    # sizes without a fixed path: REPLICATE only, width >= aligned margin + margin
    r = k // 2
    min_w = (r + 7) // 8 * 8 + r
    for idx, (w, h) in enumerate([(min_w, k - 1), (min_w + 37, k + 20), (max(min_w, 1100), k + 3)]):
        c = 1 + (k + idx) % 4  # every channel count across the sizes
        src = rng.integers(0, 256, (h, w, c), dtype=np.uint8)
        got = K.gaussian_blur(src, k, sigma, "REPLICATE")
        _, want = O.gaussian_blur(src, k, sigma, "REPLICATE")
        assert_same(got, want, "k%d s%.1f %dx%dx%d" % (k, sigma, w, h, c))


def test_gaussian_arbitrary_size_device_api_and_declines(rng):
    import torch
    src = rng.integers(0, 256, (70, 301, 4), dtype=np.uint8)
    d = K.gaussian_blur(torch.from_numpy(src).cuda(), 41, 0.0, "REPLICATE", device_api=True)
    torch.cuda.synchronize()
    assert_same(d.cpu().numpy(), O.gaussian_blur(src, 41, 0.0, "REPLICATE")[1], "device k41")
    with pytest.raises(K.KleidiCVError):
        K.gaussian_blur(src, 41, 0.0, "REFLECT")
    with pytest.raises(K.KleidiCVError):
        K.gaussian_blur(src[:, :40], 41, 0.0, "REPLICATE")


@pytest.mark.parametrize("kat", KATS["gaussian_valid_image_size"], ids=lambda k: "k%d" % k["kernel"])
def test_gaussian_kats(kat):
    src = np.array(kat["src"], np.uint8)
    assert_same(K.gaussian_blur(src, kat["kernel"], 0.0, "REVERSE"), np.array(kat["expected_sigma0"]))


def test_gaussian_padded_strides(rng):
    for c in (1, 3, 4):
        src = padded_rows(rng, 37, 61, c, 13)
        dst_buf = np.full((37, 61 * c + 7), 0xA5, np.uint8)
        dst = dst_buf[:, :61 * c].reshape(37, 61, c) if c > 1 else dst_buf[:, :61]
        K.gaussian_blur(src, 5, 0.0, "REVERSE", dst=dst)
        _, want = O.gaussian_blur(np.ascontiguousarray(src), 5, 0.0, "REVERSE")
        assert_same(dst, want)
        assert (dst_buf[:, 61 * c:] == 0xA5).all(), "row padding must be left alone"


# ---------------------------------------------------------------------------
# Separable filter
@pytest.mark.parametrize("k", [3, 5, 7])
@pytest.mark.parametrize("border", BORDERS)
def test_separable(rng, k, border):
    for (w, h) in [(k - 1, k - 1), (100, 31), (300, 17)]:
        for c in (1, 3, 4):
            for hi in (2, 9, 256):
                src = rng.integers(0, 256, (h, w, c), dtype=np.uint8)
                kx = rng.integers(0, hi, k)
                ky = rng.integers(0, hi, k)
                got = K.separable_filter_2d(src, kx, ky, border)
                _, want = O.separable_filter_2d(src, kx, ky, border)
                assert_same(got, want, "k%d %s %dx%dx%d hi%d" % (k, border, w, h, c, hi))


@pytest.mark.parametrize("kat", KATS["separable_overflow"], ids=lambda k: k["name"])
def test_separable_kats(kat):
    got = K.separable_filter_2d(np.array(kat["src"], np.uint8), kat["kernel_x"], kat["kernel_y"])
    assert_same(got, np.array(kat["expected"]))


# ---------------------------------------------------------------------------
# Median
# ---------------------------------------------------------------------------
# Blur and downsample (pyrDown)
@pytest.mark.parametrize("border", BORDERS)
def test_blur_and_downsample(rng, border):
    # test_blur_and_downsample.cpp:153-177 sizes (161 px wide), the minimum,
    # odd sizes and widths across several 128-byte output tiles
    for (w, h) in [(4, 4), (5, 5), (160, 120), (161, 121), (161, 123), (257, 33),
                   (300, 17), (1023, 19)]:
        for c in (1, 2, 3, 4):
            src = rng.integers(0, 256, (h, w, c), dtype=np.uint8)
            got = K.blur_and_downsample(src, border)
            _, want = O.blur_and_downsample(src, border)
            assert_same(got, want, "%s %dx%dx%d" % (border, w, h, c))


@pytest.mark.parametrize("kat", KATS["blur_and_downsample_minimal"])
def test_blur_and_downsample_kat(kat):
    src = np.array(kat["src"], np.uint8)
    assert_same(K.blur_and_downsample(src, kat["border"]), np.array(kat["expected"]))


def test_blur_and_downsample_padded_and_device(rng):
    import torch
    for c in (1, 3, 4):
        src = padded_rows(rng, 45, 77, c, 13)
        h, w = 45, 77
        dbuf = np.full((23, 39 * c + 7), 0xA5, np.uint8)
        dst = dbuf[:, :39 * c].reshape(23, 39, c) if c > 1 else dbuf[:, :39]
        K.blur_and_downsample(src, "REVERSE", dst=dst)
        _, want = O.blur_and_downsample(np.ascontiguousarray(src), "REVERSE")
        assert_same(dst, want, "padded c%d" % c)
        assert (dbuf[:, 39 * c:] == 0xA5).all()
        s = torch.from_numpy(np.ascontiguousarray(src)).cuda()
        d = K.blur_and_downsample(s, "REVERSE", device_api=True)
        torch.cuda.synchronize()
        assert_same(d.cpu().numpy(), want, "device c%d" % c)


@pytest.mark.parametrize("k", [3, 5, 7])
@pytest.mark.parametrize("border", BORDERS)
def test_median(rng, k, border):
    for (w, h) in [(k - 1, k - 1), (70, 23), (260, 33)]:
        for c in (1, 2, 3, 4):
            src = rng.integers(0, 256, (h, w, c), dtype=np.uint8)
            got = K.median_blur(src, k, border)
            _, want = O.median_blur(src, k, border)
            assert_same(got, want, "k%d %s %dx%dx%d" % (k, border, w, h, c))


@pytest.mark.parametrize("k", [9, 11, 15, 31, 63])
def test_median_large(rng, k):
    # window sizes the reference serves with histograms; every border
    for border in BORDERS:
        for (w, h) in [(k - 1, k - 1), (k + 45, k + 6)]:
            c = int(rng.integers(1, 5))
            src = rng.integers(0, 256, (h, w, c), dtype=np.uint8)
            got = K.median_blur(src, k, border)
            _, want = O.median_blur(src, k, border)
            assert_same(got, want, "k%d %s %dx%dx%d" % (k, border, w, h, c))


@pytest.mark.parametrize("k", [13, 21])
def test_median_hist_strips_and_moving_medians(rng, k):
    # tall enough for many strips of rows per column; a vertical ramp with
    # noise, full bright rows and dark columns moves the median far from one
    # output row to the next
    h, w = 700, 300
    ramp = np.arange(h)[:, None] * 255 // (h - 1)
    for c in (1, 3):
        src = (ramp[:, :, None] + rng.integers(-40, 41, (h, w, c))).clip(0, 255)
        src = src.astype(np.uint8)
        src[::97] = 255
        src[:, ::53] = 0
        for border in ("REPLICATE", "REFLECT", "WRAP"):
            got = K.median_blur(src, k, border)
            _, want = O.median_blur(src, k, border)
            assert_same(got, want, "k%d %s c%d" % (k, border, c))


def _median_by_box_counts(src, k, border):
    """Exact median of every k x k window from box counts of (src <= v) for
    each level v (integral images), for sizes where sorting every window
    would take minutes."""
    import pins
    r = k // 2
    p = pins.padded(src, r, border)[:, :, 0]
    h, w = src.shape[:2]
    rank = (k * k - 1) // 2
    out = np.full((h, w), 255, np.uint8)
    done = np.zeros((h, w), bool)
    for v in range(256):
        ii = np.pad(np.cumsum(np.cumsum(p <= v, axis=0), axis=1), ((1, 0), (1, 0)))
        cnt = ii[k:k + h, k:k + w] - ii[:h, k:k + w] - ii[k:k + h, :w] + ii[:h, :w]
        hit = (cnt > rank) & ~done
        out[hit] = v
        done |= hit
    return out


def test_median_large_extremes(rng):
    # the largest windows, and a low-entropy image where many values tie
    src = rng.integers(0, 256, (130, 140), dtype=np.uint8)
    assert_same(K.median_blur(src, 127, "REFLECT"), O.median_blur(src, 127, "REFLECT")[1], "k127")
    src = rng.integers(0, 256, (254, 260), dtype=np.uint8)
    assert_same(K.median_blur(src, 255, "REVERSE"), _median_by_box_counts(src, 255, "REVERSE"),
                "k255")
    src = (rng.integers(0, 3, (60, 90, 3)) * 100).astype(np.uint8)
    assert_same(K.median_blur(src, 13, "WRAP"), O.median_blur(src, 13, "WRAP")[1], "ties")


def test_median_low_entropy(rng):
    # many ties: values drawn from {0, 1, 255}
    src = rng.choice(np.array([0, 1, 255], np.uint8), (40, 90))
    for k in (3, 5, 7):
        got = K.median_blur(src, k, "REPLICATE")
        _, want = O.median_blur(src, k, "REPLICATE")
        assert_same(got, want)


# ---------------------------------------------------------------------------
# Resize
def resize_case(rng, sw, sh, dw, dh, c):
    src = rng.integers(0, 256, (sh, sw, c), dtype=np.uint8)
    got = K.resize_linear(src, dw, dh)
    err, want = O.resize_linear(src, dw, dh)
    assert err == 0
    assert_same(got, want, "%dx%d->%dx%d c%d" % (sw, sh, dw, dh, c))


@pytest.mark.parametrize("c", [1, 2, 3, 4])
def test_resize_quarter(rng, c):
    for sw, sh in [(2, 2), (34, 10), (50, 4), (98, 6), (514, 8), (640, 96), (1026, 18)]:
        resize_case(rng, sw, sh, sw // 2, sh // 2, c)
    # rows with odd pitch on both sides (unaligned vector loads and stores)
    src = padded_rows(rng, 14, 402, c, 3)
    err, want = O.resize_linear(np.ascontiguousarray(src), 201, 7)
    assert err == 0
    dst = np.full((7, 201 * c + 5), 0xEE, np.uint8)
    view = dst[:, :201 * c].reshape(7, 201, c) if c > 1 else dst[:, :201]
    K.resize_linear(src, 201, 7, dst=view)
    assert_same(view, want.reshape(view.shape), "padded quarter c%d" % c)
    assert np.all(dst[:, 201 * c:] == 0xEE)


def test_resize_exact_1ch_2x_4x(rng):
    for sw, sh in [(1, 1), (2, 3), (9, 7), (300, 20)]:
        resize_case(rng, sw, sh, 2 * sw, 2 * sh, 1)
        resize_case(rng, sw, sh, 4 * sw, 4 * sh, 1)


@pytest.mark.parametrize("c", [1, 2, 3, 4])
def test_resize_generic_sweep(rng, c):
    # upsizing (r1/r2), downsizing (r2/r3) and the edge cases the reference
    # tests: widths just above the vector minimum, pull-backs, odd sizes
    min_w = 16 // c + (1 if 16 % c else 0)
    widths = sorted({min_w, min_w + 1, 17, 23, 33, 64, 101, 149, 333, 1479})
    for sw in widths:
        if sw * c < 16:
            continue
        for dw in {sw + 1, sw * 2, (sw * 15) // 14 + 1, sw * 3 - 1, (sw * 2) // 3 + 1,
                   sw // 2 + 1, (sw + 2) // 3 + 1, sw - 1}:
            if dw <= 0:
                continue
            sh = int(rng.integers(1, 9))
            dh = int(rng.integers(1, 11))
            err, _ = O.resize_linear(np.zeros((sh, sw, c), np.uint8), dw, dh)
            if err != 0:
                with pytest.raises(K.KleidiCVError):
                    K.resize_linear(np.zeros((sh, sw, c), np.uint8), dw, dh)
                continue
            resize_case(rng, sw, sh, dw, dh, c)


def test_resize_exact_2x_4ch(rng):
    # closed-form kernel for the generic path at exactly 2x, 4 channels
    for sw, sh in [(4, 1), (4, 4), (5, 3), (17, 9), (255, 7), (256, 5),
                   (511, 2), (513, 3), (1023, 3), (1025, 4), (480, 6)]:
        resize_case(rng, sw, sh, 2 * sw, 2 * sh, 4)
    # padded rows on both sides, host and device entry points
    src = padded_rows(rng, 11, 37, 4, 5)
    err, want = O.resize_linear(np.ascontiguousarray(src), 74, 22)
    assert err == 0
    dst = np.full((22, 74 * 4 + 3), 0xEE, np.uint8)
    view = dst[:, :74 * 4].reshape(22, 74, 4)
    K.resize_linear(src, 74, 22, dst=view)
    assert_same(view, want, "padded 2x")
    assert np.all(dst[:, 74 * 4:] == 0xEE)


@pytest.mark.parametrize("kat", KATS["resize_u8_params"], ids=lambda c: "%dx%d" % (len(c["src"][0]), len(c["src"])))
def test_resize_kats(kat):
    src = np.array(kat["src"], np.uint8)
    exp = np.array(kat["expected"], np.uint8)
    assert_same(K.resize_linear(src, exp.shape[1], exp.shape[0]), exp)


@pytest.mark.parametrize("c", [1, 2, 3, 4])
def test_resize_generic_direct_reads(rng, c):
    # rows shrinking by 1.5 or more read their taps straight from global
    # memory: 8-byte windows, 16-byte windows, per-byte reads for wide tap
    # spreads, windows pulled back from the row's end, odd row pitches
    min_w = 16 // c + (1 if 16 % c else 0)
    for sw, sh, dw, dh in [(min_w, 40, max(min_w // 2, 1), 20), (min_w + 1, 9, min_w // 2 + 1, 6),
                           (257, 63, 129, 32), (301, 90, 151, 45), (515, 31, 172, 11),
                           (1031, 17, 211, 5), (777, 120, 389, 80), (2049, 33, 1025, 17)]:
        err, _ = O.resize_linear(np.zeros((sh, sw, c), np.uint8), dw, dh)
        if err != 0:
            continue
        resize_case(rng, sw, sh, dw, dh, c)
        src = padded_rows(rng, sh, sw, c, 5)
        err, want = O.resize_linear(np.ascontiguousarray(src), dw, dh)
        assert err == 0
        dst = np.full((dh, dw * c + 3), 0xEE, np.uint8)
        view = dst[:, :dw * c].reshape(dh, dw, c) if c > 1 else dst[:, :dw]
        K.resize_linear(src, dw, dh, dst=view)
        assert_same(view, want, "padded %dx%d->%dx%d c%d" % (sw, sh, dw, dh, c))
        assert np.all(dst[:, dw * c:] == 0xEE)


def test_resize_pyramid_odd_sizes(rng):
    # cfg3-style 0.5x pyramids on odd sizes leave the quarter path
    for _ in range(12):
        w, h = int(rng.integers(64, 700)), int(rng.integers(8, 80))
        c = int(rng.choice([1, 3, 4]))
        for _lvl in range(3):
            dw, dh = max(w // 2, 1), max(h // 2, 1)
            err, _ = O.resize_linear(np.zeros((h, w, c), np.uint8), dw, dh)
            if err == 0:
                resize_case(rng, w, h, dw, dh, c)
            w, h = dw, dh


# ---------------------------------------------------------------------------
# Remap
def smooth_maps(rng, dw, dh, sw, sh, amp):
    yy, xx = np.mgrid[0:dh, 0:dw].astype(np.float64)
    fx = np.zeros_like(xx)
    fy = np.zeros_like(yy)
    for _ in range(4):
        a, b, p, q = rng.uniform(0.002, 0.05, 4)
        fx += amp / 4 * np.sin(a * xx + b * yy + p * 10)
        fy += amp / 4 * np.cos(b * xx + a * yy + q * 10)
    return xx * sw / dw + fx, yy * sh / dh + fy


def s16point5_maps(mx, my):
    ix = np.floor(mx * 32).astype(np.int64)
    iy = np.floor(my * 32).astype(np.int64)
    mapxy = np.stack([ix >> 5, iy >> 5], axis=-1).astype(np.int16)
    mapfrac = ((ix & 31) | ((iy & 31) << 5)).astype(np.uint16)
    return mapxy, mapfrac


@pytest.mark.parametrize("c", [1, 4])
@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_remap_s16point5(rng, c, border):
    for (sw, sh, dw, dh) in [(31, 17, 8, 3), (200, 90, 150, 70), (64, 64, 333, 41)]:
        src = rng.integers(0, 256, (sh, sw, c), dtype=np.uint8)
        mx, my = smooth_maps(rng, dw, dh, sw, sh, 12)
        mapxy, mapfrac = s16point5_maps(mx, my)
        # sprinkle far out-of-range and extreme coordinates
        sel = rng.random((dh, dw)) < 0.05
        mapxy[sel] = rng.integers(-32768, 32767, (int(sel.sum()), 2))
        mapfrac = (mapfrac | (rng.integers(0, 64, (dh, dw)) << 10)).astype(np.uint16)
        bv = tuple(int(v) for v in rng.integers(0, 256, 4))
        got = K.remap_s16point5(src, mapxy, mapfrac, border, bv)
        _, want = O.remap_s16point5(src, mapxy, mapfrac, border, bv)
        assert_same(got, want, "%s c%d %dx%d" % (border, c, dw, dh))


@pytest.mark.parametrize("c", [1, 4])
@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_remap_s16point5_narrow_and_odd_stride(rng, c, border):
    # sources one to three pixels wide (the pair loads clamp inside a row),
    # and rows with an odd pitch (unaligned pair loads)
    cases = [(padded_rows(rng, sh, sw, c, 0), dw, dh)
             for (sw, sh, dw, dh) in [(1, 1, 8, 2), (1, 5, 9, 3), (2, 1, 8, 4), (3, 2, 13, 5)]]
    cases.append((padded_rows(rng, 23, 37, c, 3), 70, 19))
    for src, dw, dh in cases:
        sh, sw = src.shape[0], src.shape[1]
        mapxy = np.stack([rng.integers(-3, sw + 3, (dh, dw)), rng.integers(-3, sh + 3, (dh, dw))],
                         axis=-1).astype(np.int16)
        mapfrac = rng.integers(0, 1 << 16, (dh, dw)).astype(np.uint16)
        bv = tuple(int(v) for v in rng.integers(0, 256, 4))
        got = K.remap_s16point5(src, mapxy, mapfrac, border, bv)
        _, want = O.remap_s16point5(np.ascontiguousarray(src), mapxy, mapfrac, border, bv)
        assert_same(got, want, "%s c%d %dx%d -> %dx%d" % (border, c, sw, sh, dw, dh))


@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_remap_s16point5_one_channel_zooms_and_smooth_maps(rng, border):
    # Zooms by s from 1/4 to 7 (a destination tile's taps spread over 1/4 to
    # 7 times its size in the source), smooth maps reaching past the image
    # edges, a map smooth in one half and scattered in the other, odd row
    # pitches and partial tiles.
    cases = []
    for s in (0.25, 1.0, 2.5, 3.5, 3.6, 3.7, 3.8, 4.0, 7.0):
        dw, dh = 150, 37
        sw, sh = int(dw * s) + 8, int(dh * s) + 8
        yy, xx = np.mgrid[0:dh, 0:dw].astype(np.float64)
        off = rng.uniform(-6, 6, 2)
        cases.append((padded_rows(rng, sh, sw, 1, 0), xx * s + off[0], yy * s + off[1]))
    for (sw, sh, dw, dh, pad) in [(300, 120, 333, 41, 3), (131, 67, 200, 90, 0), (40, 20, 77, 13, 1)]:
        mx, my = smooth_maps(rng, dw, dh, sw, sh, 20)
        cases.append((padded_rows(rng, sh, sw, 1, pad), mx, my))
    mx, my = smooth_maps(rng, 256, 64, 256, 64, 10)
    half = np.zeros((64, 256), dtype=bool)
    half[:, 128:] = rng.random((64, 128)) < 0.02
    mx[half] = rng.uniform(-50, 300, int(half.sum()))
    my[half] = rng.uniform(-50, 120, int(half.sum()))
    cases.append((padded_rows(rng, 64, 256, 1, 0), mx, my))
    for src, mx, my in cases:
        mapxy, mapfrac = s16point5_maps(mx, my)
        bv = tuple(int(v) for v in rng.integers(0, 256, 4))
        got = K.remap_s16point5(src, mapxy, mapfrac, border, bv)
        _, want = O.remap_s16point5(np.ascontiguousarray(src), mapxy, mapfrac, border, bv)
        assert_same(got, want, "%s %s -> %s" % (border, src.shape, mapxy.shape))


@pytest.mark.parametrize("c", [1, 2, 3, 4])
@pytest.mark.parametrize("interp", ["LINEAR", "NEAREST"])
@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_remap_f32(rng, c, interp, border):
    for (sw, sh, dw, dh) in [(29, 13, 4, 2), (180, 96, 141, 77), (64, 40, 300, 9)]:
        src = rng.integers(0, 256, (sh, sw, c), dtype=np.uint8)
        mx, my = smooth_maps(rng, dw, dh, sw, sh, 20)
        mapx, mapy = mx.astype(np.float32), my.astype(np.float32)
        sel = rng.random((dh, dw)) < 0.03
        mapx[sel] = rng.choice(np.array([np.nan, np.inf, -np.inf, -1e9, 3e9, -0.5, 0.5,
                                         sw - 0.5, sw - 1 + 1e-3], np.float32), int(sel.sum()))
        bv = tuple(int(v) for v in rng.integers(0, 256, 4))
        got = K.remap_f32(src, mapx, mapy, interp, border, bv)
        _, want = O.remap_f32(src, mapx, mapy, interp, border, bv)
        assert_same(got, want, "%s %s c%d" % (interp, border, c))


# ---------------------------------------------------------------------------
# Warp perspective
def homography(rng, w, h, deg=15.0, persp=1e-4):
    a = np.deg2rad(rng.uniform(-deg, deg))
    s = rng.uniform(0.7, 1.3)
    return np.array([[s * np.cos(a), -s * np.sin(a), rng.uniform(-0.2, 0.2) * w],
                     [s * np.sin(a), s * np.cos(a), rng.uniform(-0.2, 0.2) * h],
                     [rng.uniform(-persp, persp), rng.uniform(-persp, persp), 1.0]], np.float32)


@pytest.mark.parametrize("c", [1, 2, 3, 4])
@pytest.mark.parametrize("interp", ["LINEAR", "NEAREST"])
@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_warp(rng, c, interp, border):
    for (sw, sh, dw, dh) in [(16, 4, 8, 2), (220, 130, 211, 97), (97, 61, 333, 12)]:
        src = rng.integers(0, 256, (sh, sw, c), dtype=np.uint8)
        for persp in (1e-4, 3e-3):
            T = homography(rng, sw, sh, persp=persp)
            bv = tuple(int(v) for v in rng.integers(0, 256, 4))
            got = K.warp_perspective(src, T, dw, dh, interp, border, bv)
            _, want = O.warp_perspective(src, T, dw, dh, interp, border, bv)
            assert_same(got, want, "%s %s c%d %g" % (interp, border, c, persp))


@pytest.mark.parametrize("c", [1, 2, 3, 4])
@pytest.mark.parametrize("interp", ["LINEAR", "NEAREST"])
@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_float_gathers_narrow_and_odd_stride(rng, c, interp, border):
    # sources one to three pixels wide (pair loads clamp inside a row) and
    # rows with an odd pitch (unaligned pair loads), for remap f32 and warp
    cases = [(padded_rows(rng, sh, sw, c, 0), dw, dh)
             for (sw, sh, dw, dh) in [(1, 1, 8, 2), (1, 4, 9, 3), (2, 1, 8, 4), (3, 2, 13, 5)]]
    cases.append((padded_rows(rng, 21, 35, c, 3), 67, 18))
    for src, dw, dh in cases:
        sh, sw = src.shape[0], src.shape[1]
        mapx = rng.uniform(-2.5, sw + 1.5, (dh, dw)).astype(np.float32)
        mapy = rng.uniform(-2.5, sh + 1.5, (dh, dw)).astype(np.float32)
        bv = tuple(int(v) for v in rng.integers(0, 256, 4))
        what = "%s %s c%d %dx%d" % (interp, border, c, sw, sh)
        got = K.remap_f32(src, mapx, mapy, interp, border, bv)
        _, want = O.remap_f32(np.ascontiguousarray(src), mapx, mapy, interp, border, bv)
        assert_same(got, want, "remap_f32 " + what)
        T = np.array([[sw / dw, 0.01, -0.7], [0.02, sh / dh, -0.6], [1e-3, 2e-3, 1.0]],
                     np.float32)
        got = K.warp_perspective(src, T, dw, dh, interp, border, bv)
        _, want = O.warp_perspective(np.ascontiguousarray(src), T, dw, dh, interp, border, bv)
        assert_same(got, want, "warp " + what)


def device_image_at(torch, img, lead):
    """Device copy of the (h, w[, c]) numpy view `img` (rows may be padded)
    that starts `lead` bytes into its allocation and ends at the image's last
    byte (no padding after the last row)."""
    h = img.shape[0]
    stride = img.strides[0]
    row = img.shape[1] * (img.shape[2] if img.ndim == 3 else 1)
    flat = np.zeros(lead + (h - 1) * stride + row, np.uint8)
    for y in range(h):
        flat[lead + y * stride:lead + y * stride + row] = np.ascontiguousarray(img[y]).reshape(-1)
    dev = torch.from_numpy(flat).cuda()
    shape = img.shape if img.ndim == 3 else img.shape + (1,)
    v = torch.as_strided(dev, shape, (stride, shape[2], 1), lead)
    return v if img.ndim == 3 else v[:, :, 0]


@pytest.mark.parametrize("c", [1, 2, 3, 4])
def test_gathers_at_unaligned_bases_and_image_end(rng, c):
    # Device-API gathers on sources that start 0-3 bytes past an aligned
    # address and end at the allocation's last byte, with taps crowding the
    # image's last pixels: the aligned source windows are pulled back from
    # the image's end and the lead bytes before its start are never used.
    import torch
    stream = torch.cuda.Stream()
    for (sw, sh, pad) in [(2, 1, 0), (3, 2, 1), (13, 7, 0), (50, 9, 3)]:
        img = padded_rows(rng, sh, sw, c, pad)
        host = np.ascontiguousarray(img)
        for lead in range(4):
            src = device_image_at(torch, img, lead)
            dw, dh = 37, 6
            near = rng.random((dh, dw)) < 0.6
            mx = np.where(near, rng.uniform(sw - 2.5, sw + 0.25, (dh, dw)),
                          rng.uniform(-1.5, sw + 0.5, (dh, dw))).astype(np.float32)
            my = np.where(near, rng.uniform(sh - 2.5, sh + 0.25, (dh, dw)),
                          rng.uniform(-1.5, sh + 0.5, (dh, dw))).astype(np.float32)
            bv = tuple(int(v) for v in rng.integers(0, 256, 4))
            what = "c%d %dx%d pad %d lead %d" % (c, sw, sh, pad, lead)
            for border in ("REPLICATE", "CONSTANT"):
                dst = torch.empty((dh, dw, c) if c > 1 else (dh, dw), dtype=torch.uint8,
                                  device="cuda")
                K.remap_f32(src, torch.from_numpy(mx).cuda(), torch.from_numpy(my).cuda(),
                            "LINEAR", border, bv, dst=dst, stream=stream, device_api=True)
                stream.synchronize()
                _, want = O.remap_f32(host, mx, my, "LINEAR", border, bv)
                assert_same(dst.cpu().numpy(), want, "remap_f32 %s %s" % (border, what))
                T = np.array([[1.0, 0.01, sw - dw * 0.9], [0.0, 1.0, sh - dh - 0.6],
                              [0.0, 1e-4, 1.0]], np.float32)
                K.warp_perspective(src, T, dw, dh, "LINEAR", border, bv, dst=dst, stream=stream,
                                   device_api=True)
                stream.synchronize()
                _, want = O.warp_perspective(host, T, dw, dh, "LINEAR", border, bv)
                assert_same(dst.cpu().numpy(), want, "warp %s %s" % (border, what))
                if c in (1, 4):
                    ix = np.floor(mx * 32).astype(np.int64)
                    iy = np.floor(my * 32).astype(np.int64)
                    mapxy = np.stack([ix >> 5, iy >> 5], axis=-1).astype(np.int16)
                    mapfrac = ((ix & 31) | ((iy & 31) << 5)).astype(np.uint16)
                    K.remap_s16point5(src, torch.from_numpy(mapxy).cuda(),
                                      torch.from_numpy(mapfrac).cuda(), border, bv, dst=dst,
                                      stream=stream, device_api=True)
                    stream.synchronize()
                    _, want = O.remap_s16point5(host, mapxy, mapfrac, border, bv)
                    assert_same(dst.cpu().numpy(), want, "remap_s16 %s %s" % (border, what))


def test_warp_zero_weight_row(rng):
    # a homography whose weight crosses 0 inside the image takes the checked path
    src = rng.integers(0, 256, (40, 60), dtype=np.uint8)
    T = np.array([[1, 0, 0], [0, 1, 0], [0.02, 0, -0.5]], np.float32)
    for interp in ("LINEAR", "NEAREST"):
        for border in ("REPLICATE", "CONSTANT"):
            got = K.warp_perspective(src, T, 64, 40, interp, border, (5,))
            _, want = O.warp_perspective(src, T, 64, 40, interp, border, (5,))
            assert_same(got, want)


def test_warp_kat():
    kat = KATS["warp_simple_resize"]
    src = np.array(kat["src"], np.uint8)
    exp = np.array(kat["expected"], np.uint8)
    assert_same(K.warp_perspective(src, kat["transform"], exp.shape[1], exp.shape[0]), exp)


# ---------------------------------------------------------------------------
# NV12
@pytest.mark.parametrize("fmt", [0x02, 0x22, 0x42, 0x62, 0x12, 0x32, 0x52, 0x72])
def test_nv12(rng, fmt):
    for (w, h) in [(1, 1), (3, 5), (16, 2), (101, 37), (640, 33)]:
        y = rng.integers(0, 256, (h, w), dtype=np.uint8)
        uv = rng.integers(0, 256, ((h + 1) // 2, (w + 1) // 2 * 2), dtype=np.uint8)
        got = K.yuv_sp_to_rgb(y, uv, fmt)
        _, want = O.yuv_sp_to_rgb(y, uv, fmt)
        assert_same(got, want, "fmt %#x %dx%d" % (fmt, w, h))


# ---------------------------------------------------------------------------
# Device-resident API on a torch stream gives the same bytes as the host API.
# ---------------------------------------------------------------------------
# One-buffer YUV layouts (test_yuv420p_to_rgb.cpp:58-69, test_yuv422_to_rgb.cpp,
# test_yuv444_to_rgb.cpp sizes, plus widths across several thread blocks)
FMT_420P = [0x03, 0x23, 0x43, 0x63, 0x13, 0x33, 0x53, 0x73]
FMT_422 = [b | f for b in (0x04, 0x84, 0x14) for f in (0x00, 0x20, 0x40, 0x60)]
FMT_444 = [0x01, 0x21, 0x41, 0x61]


@pytest.mark.parametrize("fmt", FMT_420P)
def test_yuv420p(rng, fmt):
    for w in (2, 4, 6, 18, 33, 2049, 4100):
        for h in (1, 2, 3, 5, 11, 16, 22):
            src = rng.integers(0, 256, ((3 * h + 1) // 2, w + 4), dtype=np.uint8)
            got = K.yuv_to_rgb(src, w, h, fmt)
            _, want = O.yuv_to_rgb(src, w, h, fmt)
            assert_same(got, want, "fmt %#x %dx%d" % (fmt, w, h))


@pytest.mark.parametrize("fmt", FMT_422)
def test_yuv422(rng, fmt):
    for w in (2, 4, 18, 64, 2050, 4098):
        for h in (1, 3, 10):
            src = rng.integers(0, 256, (h, 2 * w + 6), dtype=np.uint8)
            got = K.yuv_to_rgb(src, w, h, fmt)
            _, want = O.yuv_to_rgb(src, w, h, fmt)
            assert_same(got, want, "fmt %#x %dx%d" % (fmt, w, h))


@pytest.mark.parametrize("fmt", FMT_444)
def test_yuv444(rng, fmt):
    for w in (1, 2, 5, 17, 2049, 4100):
        for h in (1, 4, 9):
            for pad in (0, 5):
                src = rng.integers(0, 256, (h, 3 * w + pad), dtype=np.uint8)
                got = K.yuv_to_rgb(src, w, h, fmt)
                _, want = O.yuv_to_rgb(src, w, h, fmt)
                assert_same(got, want, "fmt %#x %dx%d pad %d" % (fmt, w, h, pad))


def test_yuv_layouts_device_api_and_padded_dst(rng):
    import torch
    for fmt, rows, row_bytes in ((0x63, lambda h: (3 * h + 1) // 2, lambda w: w),
                                 (0x84, lambda h: h, lambda w: 2 * w),
                                 (0x21, lambda h: h, lambda w: 3 * w)):
        w, h = 1030, 26
        oc = 4 if fmt & 0x40 else 3
        src = rng.integers(0, 256, (rows(h), row_bytes(w) + 16), dtype=np.uint8)
        _, want = O.yuv_to_rgb(src, w, h, fmt)
        dbuf = np.full((h, w * oc + 9), 0x5A, np.uint8)
        dst = dbuf[:, :w * oc].reshape(h, w, oc)
        K.yuv_to_rgb(src, w, h, fmt, dst=dst)
        assert_same(dst, want, "padded %#x" % fmt)
        assert (dbuf[:, w * oc:] == 0x5A).all()
        d = K.yuv_to_rgb(torch.from_numpy(src).cuda(), w, h, fmt, device_api=True)
        torch.cuda.synchronize()
        assert_same(d.cpu().numpy(), want, "device %#x" % fmt)


# ---------------------------------------------------------------------------
# RGB/BGR(A) to YUV (test_rgb_to_yuv420p.cpp:80-86 sizes, odd sizes, widths
# across thread blocks)
SRC_FLAGS = [0x00, 0x20, 0x40, 0x60]


@pytest.mark.parametrize("flags", SRC_FLAGS)
def test_rgb_to_yuv420(rng, flags):
    scn = 4 if flags & 0x40 else 3
    for w in (2, 18, 32, 33, 770, 2051):
        for h in (1, 2, 11, 16, 21):
            src = rng.integers(0, 256, (h, w, scn), dtype=np.uint8)
            for vu in (0, 0x10):
                fmt = 0x03 | flags | vu
                _, want = O.rgb_to_yuv(src, fmt, dst_stride=w + 2)
                dst = np.zeros(((3 * h + 1) // 2, w + 2), np.uint8)
                K.rgb_to_yuv(src, fmt, dst=dst)
                assert_same(dst[:, :w], want[:, :w], "planar %#x %dx%d" % (fmt, w, h))
                fmt = 0x02 | flags | vu
                _, wy, wuv = O.rgb_to_yuv_semiplanar(src, fmt)
                gy, guv = K.rgb_to_yuv_semiplanar(src, fmt)
                assert_same(gy, wy, "nv luma %#x %dx%d" % (fmt, w, h))
                assert_same(guv, wuv, "nv chroma %#x %dx%d" % (fmt, w, h))


@pytest.mark.parametrize("flags", SRC_FLAGS)
def test_rgb_to_yuv422_444(rng, flags):
    scn = 4 if flags & 0x40 else 3
    for w in (2, 6, 18, 64, 2050):
        for h in (1, 3, 10):
            src = rng.integers(0, 256, (h, w, scn), dtype=np.uint8)
            for base in (0x04, 0x84, 0x14, 0x01):
                fmt = base | flags
                _, want = O.rgb_to_yuv(src, fmt)
                assert_same(K.rgb_to_yuv(src, fmt), want, "%#x %dx%d" % (fmt, w, h))
    src = rng.integers(0, 256, (4, 9, scn), dtype=np.uint8)  # odd width: 4:4:4 only
    assert_same(K.rgb_to_yuv(src, 0x01 | flags), O.rgb_to_yuv(src, 0x01 | flags)[1], "444 odd")


def test_rgb_to_yuv_device_api(rng):
    import torch
    src = rng.integers(0, 256, (26, 1030, 4), dtype=np.uint8)
    s = torch.from_numpy(src).cuda()
    for fmt in (0x43, 0x64, 0x41):
        d = K.rgb_to_yuv(s, fmt, device_api=True)
        torch.cuda.synchronize()
        assert_same(d.cpu().numpy(), O.rgb_to_yuv(src, fmt)[1], "device %#x" % fmt)
    y, uv = K.rgb_to_yuv_semiplanar(s, 0x52, device_api=True)
    torch.cuda.synchronize()
    _, wy, wuv = O.rgb_to_yuv_semiplanar(src, 0x52)
    assert_same(y.cpu().numpy(), wy, "device nv21 luma")
    assert_same(uv.cpu().numpy(), wuv, "device nv21 chroma")


def test_device_api_matches_host_api(rng):
    import torch
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    src = rng.integers(0, 256, (120, 250, 3), dtype=np.uint8)
    d_src = torch.from_numpy(src).to(dev)
    with torch.cuda.stream(stream):
        outs = {
            "gauss": K.gaussian_blur(d_src, 5, 0.0, "REVERSE", stream=stream, device_api=True),
            "median": K.median_blur(d_src, 5, "REPLICATE", stream=stream, device_api=True),
            "resize": K.resize_linear(d_src, 125, 60, stream=stream, device_api=True),
            "resize_up": K.resize_linear(d_src, 333, 200, stream=stream, device_api=True),
        }
    stream.synchronize()
    assert_same(outs["gauss"].cpu().numpy(), K.gaussian_blur(src, 5, 0.0, "REVERSE"))
    assert_same(outs["median"].cpu().numpy(), K.median_blur(src, 5, "REPLICATE"))
    assert_same(outs["resize"].cpu().numpy(), K.resize_linear(src, 125, 60))
    assert_same(outs["resize_up"].cpu().numpy(), K.resize_linear(src, 333, 200))


# ---------------------------------------------------------------------------
# The warp's refined reciprocal equals the IEEE quotient over every float of
# its range (the library's own self-check kernel; about 0.1 s).
def test_refined_reciprocal_is_correctly_rounded():
    import ctypes as C
    fn = K.lib().kcv_selfcheck_reciprocal
    fn.argtypes = [C.POINTER(C.c_ulonglong)]
    fn.restype = C.c_int
    bad = C.c_ulonglong(123)
    assert fn(C.byref(bad)) == 0
    assert bad.value == 0


# The lerp rounding of the float gathers (a hardware conversion) equals
# round-half-away over every float in [0, 256] (self-check kernel).
def test_lerp_rounding_is_half_away():
    import ctypes as C
    fn = K.lib().kcv_selfcheck_rounding
    fn.argtypes = [C.POINTER(C.c_ulonglong)]
    fn.restype = C.c_int
    bad = C.c_ulonglong(123)
    assert fn(C.byref(bad)) == 0
    assert bad.value == 0
