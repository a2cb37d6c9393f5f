"""Pins the CPU oracle (test infrastructure) before anything is compared
against it: the reference tests' known-answer tables, the generic resize
generator's direct formula from the reference's unit test, and independent
numpy formulations (numpy.pad borders, window sums, numpy.median, double
precision geometry).  CPU only."""

import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import pins

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
BORDERS = ["REPLICATE", "REFLECT", "WRAP", "REVERSE"]


def test_border_index_matches_numpy_pad():
    for border, mode in pins.PAD_MODE.items():
        for n in (3, 4, 7):
            r = n - 1  # a k x k filter needs n >= k-1, so overhangs stay below n
            idx = np.pad(np.arange(n), (r, r), mode=mode)
            got = [O.lib().orc_border_index(i, n, O.BORDER[border]) for i in range(-r, n + r)]
            assert list(idx) == got, (border, n)


@pytest.mark.parametrize("kat", KATS["gaussian_valid_image_size"], ids=lambda k: "k%d" % k["kernel"])
def test_gaussian_kats(kat):
    src = np.array(kat["src"], np.uint8)
    err, out = O.gaussian_blur(src, kat["kernel"], 0.0, kat["border"])
    assert err == 0 and (out == np.array(kat["expected_sigma0"])).all()
    err, out = O.gaussian_blur(src, kat["kernel"], 2.25, kat["border"])
    assert err == 0
    assert np.abs(out.astype(int) - np.array(kat["expected_sigma2_25"])).max() <= kat["sigma2_25_tolerance"]


@pytest.mark.parametrize("k", [3, 5, 7, 9])
@pytest.mark.parametrize("border", BORDERS)
@pytest.mark.parametrize("channels", [1, 3, 4])
def test_gaussian_binomial_vs_padded_convolution(rng, k, border, channels):
    for h, w in [(k - 1, k - 1), (11, 23), (17, 9)]:
        src = rng.integers(0, 256, (h, w, channels), dtype=np.uint8)
        err, out = O.gaussian_blur(src, k, 0.0, border)
        assert err == 0
        np.testing.assert_array_equal(out, pins.binomial_blur(src, k, border))


def test_gaussian_sigma_rounds_twice(rng):
    # The sigma path rounds the vertical pass to u8 before the horizontal
    # one (gaussian_blur_fixed_neon.cpp:461-529): check on a case where a
    # single rounding would differ.
    half = np.zeros(3, np.uint8)
    assert O.lib().orc_gaussian_half_kernel(half.ctypes.data, 3, 1.0)
    w = np.concatenate([half, half[-2::-1]]).astype(np.int64)
    assert w.sum() == 256
    src = rng.integers(0, 256, (12, 12), dtype=np.uint8)
    err, out = O.gaussian_blur(src, 5, 1.0, "REPLICATE")
    p = np.pad(src, 2, mode="edge").astype(np.int64)
    vert = sum(w[i] * p[i:i + 12] for i in range(5))
    vert = (vert + 128) >> 8
    hor = sum(w[j] * vert[:, j:j + 12] for j in range(5))
    np.testing.assert_array_equal(out, ((hor + 128) >> 8).astype(np.uint8))


@pytest.mark.parametrize("k,sigma", [(11, 0.0), (31, 4.0), (63, 0.0)])
def test_gaussian_arbitrary_size_rounds_twice(rng, k, sigma):
    # sizes without a fixed path (gaussian_blur_arbitrary_neon.cpp): REPLICATE,
    # vertical pass rounded to u8, then the horizontal pass
    r = k // 2
    half = np.zeros(r + 1, np.uint8)
    assert O.lib().orc_gaussian_half_kernel(half.ctypes.data, r + 1, C.c_float(sigma))
    w = np.concatenate([half, half[-2::-1]]).astype(np.int64)
    assert w.sum() == 256 and len(w) == k
    h, wd, c = k + 3, (r + 7) // 8 * 8 + r + 5, 3
    src = rng.integers(0, 256, (h, wd, c), dtype=np.uint8)
    err, out = O.gaussian_blur(src, k, sigma, "REPLICATE")
    assert err == 0
    p = np.pad(src, ((r, r), (r, r), (0, 0)), mode="edge").astype(np.int64)
    vert = (sum(w[i] * p[i:i + h] for i in range(k)) + 128) >> 8
    hor = (sum(w[j] * vert[:, j:j + wd] for j in range(k)) + 128) >> 8
    np.testing.assert_array_equal(out, hor.astype(np.uint8))
    # only REPLICATE, and only images at least aligned_margin + margin wide
    assert O.gaussian_blur(src, k, sigma, "REFLECT")[0] == 1
    narrow = src[:, :(r + 7) // 8 * 8 + r - 1]
    assert O.gaussian_blur(narrow, k, sigma, "REPLICATE")[0] == 1


@pytest.mark.parametrize("kat", KATS["blur_and_downsample_minimal"])
def test_blur_and_downsample_kat(kat):
    err, out = O.blur_and_downsample(np.array(kat["src"], np.uint8), kat["border"])
    assert err == 0
    np.testing.assert_array_equal(out, np.array(kat["expected"]))


@pytest.mark.parametrize("border", BORDERS)
@pytest.mark.parametrize("channels", [1, 2, 3, 4])
def test_blur_and_downsample_vs_padded_convolution(rng, border, channels):
    for h, w in [(4, 4), (5, 5), (7, 9), (12, 161), (9, 30)]:
        src = rng.integers(0, 256, (h, w, channels), dtype=np.uint8)
        err, out = O.blur_and_downsample(src, border)
        assert err == 0
        np.testing.assert_array_equal(out, pins.blur_and_downsample(src, border))
        # the even-position samples of the full Gaussian 5x5
        _, full = O.gaussian_blur(src, 5, 0.0, border)
        np.testing.assert_array_equal(out, full[::2, ::2])


def test_blur_and_downsample_stripes_and_errors(rng):
    src = rng.integers(0, 256, (23, 17, 3), dtype=np.uint8)
    _, whole = O.blur_and_downsample(src, "WRAP")
    parts = np.zeros_like(whole)
    for y0, y1 in [(0, 5), (5, 6), (6, 15), (15, 23)]:
        err, part = O.blur_and_downsample(src, "WRAP", (y0, y1))
        assert err == 0
        rows = slice((y0 + 1) // 2, (y1 + 1) // 2)
        parts[rows] = part[rows]
    np.testing.assert_array_equal(parts, whole)
    # test_blur_and_downsample.cpp:208-287: borders, undersize, channel count
    for border in ("CONSTANT", "TRANSPARENT", "NONE"):
        assert O.blur_and_downsample(src, border)[0] == 1
    assert O.blur_and_downsample(src[:3], "REPLICATE")[0] == 1
    assert O.blur_and_downsample(src[:, :3], "REPLICATE")[0] == 1


@pytest.mark.parametrize("kat", KATS["separable_overflow"], ids=lambda k: k["name"])
def test_separable_kats(kat):
    err, out = O.separable_filter_2d(np.array(kat["src"], np.uint8), kat["kernel_x"],
                                     kat["kernel_y"], kat["border"])
    assert err == 0 and (out == np.array(kat["expected"])).all()


@pytest.mark.parametrize("k", [3, 5, 7])
@pytest.mark.parametrize("border", BORDERS)
def test_separable_vs_two_stage_saturation(rng, k, border):
    for taps_hi in (3, 40, 256):
        for channels in (1, 2, 4):
            src = rng.integers(0, 256, (13, 19, channels), dtype=np.uint8)
            kx = rng.integers(0, taps_hi, k)
            ky = rng.integers(0, taps_hi, k)
            err, out = O.separable_filter_2d(src, kx, ky, border)
            assert err == 0
            np.testing.assert_array_equal(out, pins.separable_two_stage(src, kx, ky, border))


@pytest.mark.parametrize("k", [3, 5, 7])
@pytest.mark.parametrize("border", BORDERS)
def test_median_vs_numpy(rng, k, border):
    for h, w, c in [(k - 1, k - 1, 1), (9, 21, 3), (14, 6, 4)]:
        src = rng.integers(0, 256, (h, w, c), dtype=np.uint8)
        err, out = O.median_blur(src, k, border)
        assert err == 0
        np.testing.assert_array_equal(out, pins.median(src, k, border))


@pytest.mark.parametrize("case", KATS["resize_u8_params"], ids=lambda c: "%dx%d" % (len(c["src"][0]), len(c["src"])))
def test_resize_param_tables(case):
    src = np.array(case["src"], np.uint8)
    exp = np.array(case["expected"], np.uint8)
    err, out = O.resize_linear(src, exp.shape[1], exp.shape[0])
    assert err == 0
    np.testing.assert_array_equal(out, exp)


@pytest.mark.parametrize("g", KATS["resize_generator_sizes"],
                         ids=lambda g: "c%d_%d_%d" % (g["channels"], g["src_width"], g["dst_width"]))
def test_resize_generator_vs_direct_formula(g):
    C, sw, dw, tol = g["channels"], g["src_width"], g["dst_width"], g["xfrac_tolerance"]
    # the unit test's own choice of generator variant
    R = 2 if (dw > sw and dw * 14 <= sw * 15) else (sw // dw) + 1
    up = dw >= sw
    left, right, frac, _ = O.resize_columns(sw, dw, C, R, up)
    l_ref, r_ref, f_ref = pins.resize_generator_direct(C, sw, dw, R, up)
    np.testing.assert_array_equal(left, l_ref)
    np.testing.assert_array_equal(right, r_ref)
    assert np.abs(frac.astype(int) - f_ref).max() <= tol


@pytest.mark.parametrize("channels", [1, 2, 3, 4])
def test_resize_quarter_closed_form(rng, channels):
    src = rng.integers(0, 256, (18, 26, channels), dtype=np.uint8)
    err, out = O.resize_linear(src, 13, 9)
    s = src.astype(np.int64)
    ref = ((s[0::2, 0::2] + s[0::2, 1::2] + s[1::2, 0::2] + s[1::2, 1::2] + 2) >> 2)
    assert err == 0
    np.testing.assert_array_equal(out, ref.astype(np.uint8))


@pytest.mark.parametrize("channels", [2, 3, 4])
@pytest.mark.parametrize("h,w", [(6, 24), (1, 17), (3, 255), (2, 257), (4, 600), (9, 13)])
def test_resize_exact_2x_generic_equals_direct_lerp(rng, channels, h, w):
    # Exact 2x upsizing goes through the generic path for C > 1; every
    # fraction is exact (64 or 192 in 8 bits), so the two-stage lerp on the
    # textbook coordinates must reproduce it, across vector tails and the
    # generator's recalibration points.
    if w * channels < 16:
        pytest.skip("below the generic path's minimum width")
    src = rng.integers(0, 256, (h, w, channels), dtype=np.uint8)
    err, out = O.resize_linear(src, 2 * w, 2 * h)
    assert err == 0

    def coords(n_dst, n_src):
        s = (np.arange(n_dst) + 0.5) / 2 - 0.5
        i0 = np.floor(s).astype(int)
        f = np.round((s - np.floor(s)) * 256).astype(int)
        return np.clip(i0, 0, n_src - 1), np.clip(i0 + 1, 0, n_src - 1), f

    x0, x1, fx = coords(2 * w, w)
    y0, y1, fy = coords(2 * h, h)
    s = src.astype(np.int64)
    lerp = lambda a, b, t: (256 * a + (b - a) * t + 128) >> 8
    left = lerp(s[y0][:, x0], s[y1][:, x0], fy[:, None, None])
    right = lerp(s[y0][:, x1], s[y1][:, x1], fy[:, None, None])
    ref = lerp(left, right, fx[None, :, None])
    np.testing.assert_array_equal(out, ref.astype(np.uint8))


@pytest.mark.parametrize("channels", [1, 4])
@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_remap_s16point5_vs_reference_closed_form(rng, channels, border):
    h, w, dh, dw = 20, 31, 17, 29
    src = rng.integers(0, 256, (h, w, channels), dtype=np.uint8)
    mapxy = np.stack([rng.integers(-5, w + 5, (dh, dw)), rng.integers(-5, h + 5, (dh, dw))],
                     axis=-1).astype(np.int16)
    mapfrac = rng.integers(0, 1024, (dh, dw)).astype(np.uint16)
    bv = (7, 99, 200, 3)
    err, out = O.remap_s16point5(src, mapxy, mapfrac, border, bv)
    assert err == 0
    np.testing.assert_array_equal(out, pins.remap_s16point5(src, mapxy, mapfrac, border, bv))


@pytest.mark.parametrize("interp", ["LINEAR", "NEAREST"])
@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
@pytest.mark.parametrize("channels", [1, 2, 4])
def test_remap_f32_within_one_of_double(rng, interp, border, channels):
    h, w, dh, dw = 23, 37, 19, 33
    src = rng.integers(0, 256, (h, w, channels), dtype=np.uint8)
    mapx = rng.uniform(-3, w + 3, (dh, dw)).astype(np.float32)
    mapy = rng.uniform(-3, h + 3, (dh, dw)).astype(np.float32)
    bv = (11, 22, 33, 44)
    err, out = O.remap_f32(src, mapx, mapy, interp, border, bv)
    assert err == 0
    ref = pins.remap_f32_double(src, mapx, mapy, interp, border, bv)
    assert np.abs(out.astype(int) - ref.astype(int)).max() <= 1
    # integer coordinates copy pixels exactly
    ix = rng.integers(0, w, (dh, dw)).astype(np.float32)
    iy = rng.integers(0, h, (dh, dw)).astype(np.float32)
    err, out = O.remap_f32(src, ix, iy, interp, border, bv)
    np.testing.assert_array_equal(out, src[iy.astype(int), ix.astype(int)])


def test_remap_f32_per_plane_invariant(rng):
    src = rng.integers(0, 256, (21, 30, 4), dtype=np.uint8)
    mapx = rng.uniform(-4, 34, (15, 26)).astype(np.float32)
    mapy = rng.uniform(-4, 25, (15, 26)).astype(np.float32)
    bv = (1, 2, 3, 4)
    for interp in ("LINEAR", "NEAREST"):
        for border in ("REPLICATE", "CONSTANT"):
            _, four = O.remap_f32(src, mapx, mapy, interp, border, bv)
            for c in range(4):
                _, one = O.remap_f32(np.ascontiguousarray(src[..., c]), mapx, mapy, interp,
                                     border, (bv[c],))
                np.testing.assert_array_equal(four[..., c], one)


def test_warp_simple_resize_kat():
    kat = KATS["warp_simple_resize"]
    src = np.array(kat["src"], np.uint8)
    exp = np.array(kat["expected"], np.uint8)
    err, out = O.warp_perspective(src, kat["transform"], exp.shape[1], exp.shape[0],
                                  kat["interpolation"], kat["border"])
    assert err == 0
    np.testing.assert_array_equal(out, exp)


def random_homography(rng, w, h, max_deg=15.0, persp=1e-4):
    a = np.deg2rad(rng.uniform(-max_deg, max_deg))
    s = rng.uniform(0.8, 1.2)
    T = np.array([[s * np.cos(a), -s * np.sin(a), rng.uniform(-0.1, 0.1) * w],
                  [s * np.sin(a), s * np.cos(a), rng.uniform(-0.1, 0.1) * h],
                  [rng.uniform(-persp, persp), rng.uniform(-persp, persp), 1.0]])
    return T.astype(np.float32)


@pytest.mark.parametrize("interp", ["LINEAR", "NEAREST"])
@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_warp_within_one_of_double(rng, interp, border):
    src = rng.integers(0, 256, (40, 56), dtype=np.uint8)
    T = random_homography(rng, 56, 40, persp=3e-3)
    err, out = O.warp_perspective(src, T, 48, 36, interp, border, (77,))
    assert err == 0
    ref = pins.warp_double(src, T, 48, 36, interp, border, (77,))
    diff = np.abs(out.astype(int) - ref.astype(int))
    if interp == "LINEAR":
        assert diff.max() <= 1
    else:
        # float vs double coordinates may straddle a rounding tie
        assert (diff > 1).mean() < 0.01


def test_warp_per_plane_invariant(rng):
    src = rng.integers(0, 256, (30, 40, 4), dtype=np.uint8)
    T = random_homography(rng, 40, 30, persp=2e-3)
    bv = (9, 8, 7, 6)
    for interp in ("LINEAR", "NEAREST"):
        for border in ("REPLICATE", "CONSTANT"):
            _, four = O.warp_perspective(src, T, 36, 28, interp, border, bv)
            for c in range(4):
                _, one = O.warp_perspective(np.ascontiguousarray(src[..., c]), T, 36, 28,
                                            interp, border, (bv[c],))
                np.testing.assert_array_equal(four[..., c], one)


@pytest.mark.parametrize("fmt", [0x02, 0x22, 0x42, 0x62, 0x12, 0x32, 0x52, 0x72])
def test_nv12_fixed_inputs_and_bt601(fmt):
    fx = KATS["nv12_fixed_inputs"]
    y = np.array(fx["y"], np.uint8)
    uv = np.array(fx["uv"], np.uint8)
    err, out = O.yuv_sp_to_rgb(y, uv, fmt)
    assert err == 0
    ref = pins.nv12_float(y, uv, vu=bool(fmt & 0x10), bgr=bool(fmt & 0x20), alpha=bool(fmt & 0x40))
    assert np.abs(out.astype(int) - ref.astype(int)).max() <= 1


def test_nv12_gray_axis_exact(rng):
    # U = V = 128 removes the chroma terms: R = G = B = round(1220542 Y' / 2^20)
    y = rng.integers(0, 256, (6, 10), dtype=np.uint8)
    uv = np.full((3, 10), 128, np.uint8)
    _, out = O.yuv_sp_to_rgb(y, uv, 0x02)
    g = (1220542 * np.maximum(y.astype(np.int64) - 16, 0) + (1 << 19)) >> 20
    g = np.clip(g, 0, 255)
    for c in range(3):
        np.testing.assert_array_equal(out[..., c], g)


# ---------------------------------------------------------------------------
# One-buffer YUV layouts
def _nv12_ref(yplane, uplane, vplane, fmt_flags):
    h, w = yplane.shape
    uv = np.stack([uplane, vplane], axis=-1).reshape(uplane.shape[0], -1)
    err, out = O.yuv_sp_to_rgb(yplane, uv, 0x02 | fmt_flags)
    assert err == 0
    return out


@pytest.mark.parametrize("h,w", [(2, 2), (4, 6), (6, 18), (16, 34), (10, 8)])
@pytest.mark.parametrize("flags", [0x00, 0x20, 0x40, 0x60])
def test_i420_yv12_equal_nv12_on_the_same_planes(rng, h, w, flags):
    for yv12 in (False, True):
        yplane = rng.integers(0, 256, (h, w), dtype=np.uint8)
        up = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
        vp = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
        buf = pins.i420_buffer(yplane, up, vp, w + 4, yv12)
        err, got = O.yuv_to_rgb(buf, w, h, 0x03 | flags | (0x10 if yv12 else 0))
        assert err == 0
        np.testing.assert_array_equal(got, _nv12_ref(yplane, up, vp, flags))


@pytest.mark.parametrize("order", ["YUYV", "UYVY", "YVYU"])
@pytest.mark.parametrize("flags", [0x00, 0x20, 0x40, 0x60])
def test_yuv422_equals_nv12_with_chroma_on_both_rows(rng, order, flags):
    h, w = 6, 20
    yplane = rng.integers(0, 256, (h, w), dtype=np.uint8)
    up = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
    vp = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
    src = np.zeros((h, 2 * w + 6), np.uint8)
    for y in range(h):
        for p in range(w // 2):
            y0, y1 = yplane[y, 2 * p], yplane[y, 2 * p + 1]
            u, v = up[y // 2, p], vp[y // 2, p]
            src[y, 4 * p:4 * p + 4] = {"YUYV": (y0, u, y1, v), "UYVY": (u, y0, v, y1),
                                       "YVYU": (y0, v, y1, u)}[order]
    base = {"YUYV": 0x04, "UYVY": 0x84, "YVYU": 0x14}[order]
    err, got = O.yuv_to_rgb(src, w, h, base | flags)
    assert err == 0
    np.testing.assert_array_equal(got, _nv12_ref(yplane, up, vp, flags))


@pytest.mark.parametrize("flags", [0x00, 0x20, 0x40, 0x60])
def test_yuv444_vector_form(rng, flags):
    h, w = 7, 33
    src = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    # the reference's fixed rows (test_yuv444_to_rgb.cpp:45-48) in front
    fixed = np.array([[10, 100, 130, 20, 100, 130, 255, 255, 255, 199, 255, 255],
                      [1, 100, 130, 120, 100, 130, 0, 255, 255, 17, 255, 255],
                      [2, 0, 1, 3, 0, 1, 240, 3, 4, 228, 3, 4],
                      [7, 7, 8, 11, 7, 8, 128, 9, 10, 129, 9, 10]], np.uint8)
    src[:4, :4] = fixed.reshape(4, 4, 3)
    err, got = O.yuv_to_rgb(src.reshape(h, -1), w, h, 0x01 | flags)
    assert err == 0
    np.testing.assert_array_equal(got, pins.yuv444_vector_form(src, bool(flags & 0x20),
                                                               bool(flags & 0x40)))


def test_yuv_routing_errors():
    buf = np.zeros((16, 64), np.uint8)
    assert O.yuv_to_rgb(buf, 10, 10, 0x83)[0] == 1     # 4:2:0 with chroma-first
    assert O.yuv_to_rgb(buf, 10, 10, 0x11)[0] == 1     # 4:4:4 with VU
    assert O.yuv_to_rgb(buf, 9, 4, 0x04)[0] == 1       # 4:2:2 odd width
    assert O.yuv_to_rgb(buf, 10, 4, 0x94)[0] == 1      # VU and chroma-first
    assert O.yuv_to_rgb(buf, 10, 4, 0x02)[0] == 1      # NV12 is not one-buffer
    assert O.yuv_to_rgb(buf, 10, 4, 0x00)[0] == 1
    assert O.yuv_to_rgb(buf, 0, 1, 0x03)[0] == 0


# ---------------------------------------------------------------------------
# RGB to YUV
@pytest.mark.parametrize("h,w", [(16, 18), (11, 32), (10, 770), (2, 2)])
@pytest.mark.parametrize("flags", [0x00, 0x20, 0x40, 0x60])
def test_rgb_to_i420_yv12_and_nv12_vs_reference_formula(rng, h, w, flags):
    scn = 4 if flags & 0x40 else 3
    src = rng.integers(0, 256, (h, w, scn), dtype=np.uint8)
    rgb = src[..., [2, 1, 0]] if flags & 0x20 else src[..., :3]
    y, u, v = pins.rgb_to_yuv420_reference_formula(rgb)
    for vu in (False, True):
        err, buf = O.rgb_to_yuv(src, 0x03 | flags | (0x10 if vu else 0), dst_stride=w + 2)
        assert err == 0
        np.testing.assert_array_equal(buf[:h, :w], y)
        if h % 2 == 0:   # pins.i420_buffer lays out the even-height chroma area
            want = pins.i420_buffer(y, u, v, w + 2, vu)
            np.testing.assert_array_equal(buf[h:, :w], want[h:, :w])
        err, yy, uv = O.rgb_to_yuv_semiplanar(src, 0x02 | flags | (0x10 if vu else 0))
        assert err == 0
        np.testing.assert_array_equal(yy, y)
        np.testing.assert_array_equal(uv[:, 0::2], v if vu else u)
        np.testing.assert_array_equal(uv[:, 1::2], u if vu else v)


def test_rgb_to_yuv_gray_axis():
    # gray maps to U = V = 128 in every layout; 4:4:4 keeps Y = gray level,
    # the limited-range forms put 0 at 16 and 255 at 235
    levels = np.arange(256, dtype=np.uint8)
    src = np.repeat(np.repeat(levels[None, :, None], 2, axis=0), 3, axis=2)
    err, b444 = O.rgb_to_yuv(src, 0x01)
    assert err == 0
    t = b444[0].reshape(-1, 3)
    assert (t[:, 0] == levels).all() and (t[:, 1] == 128).all() and (t[:, 2] == 128).all()
    err, y, uv = O.rgb_to_yuv_semiplanar(src, 0x02)
    assert err == 0 and (uv == 128).all() and y[0, 0] == 16 and y[0, 255] == 235
    assert (np.diff(y[0].astype(int)) >= 0).all()
    err, b422 = O.rgb_to_yuv(src, 0x04)
    assert err == 0
    q = b422[0].reshape(-1, 4)
    assert (q[:, 1] == 128).all() and (q[:, 3] == 128).all()
    # the same limited-range luma, up to the 14- vs 20-bit weights' rounding
    assert np.abs(q[:, 0].astype(int) - y[0, 0::2].astype(int)).max() <= 1
    assert q[0, 0] == 16 and q[-1, 2] == 235


@pytest.mark.parametrize("fmt", [0x01, 0x21, 0x41, 0x61])
def test_rgb_to_yuv444_round_trip(rng, fmt):
    # 4:4:4 there and back stays within 2 of the source for colours whose
    # chroma does not clip (V = 0.877 (R - Y) leaves 8 bits for saturated
    # reds and cyans, so channels are kept within 48..208)
    scn = 4 if fmt & 0x40 else 3
    src = rng.integers(48, 209, (9, 31, scn), dtype=np.uint8)
    err, buf = O.rgb_to_yuv(src, fmt)
    assert err == 0
    err, back = O.yuv_to_rgb(buf, 31, 9, 0x01 | (fmt & 0x20))
    assert err == 0
    d = np.abs(back.astype(int) - src[..., :3].astype(int))
    assert d.max() <= 2


@pytest.mark.parametrize("order,base", [("YUYV", 0x04), ("UYVY", 0x84), ("YVYU", 0x14)])
def test_rgb_to_yuv422_pairs(rng, order, base):
    # luma equals the 4:2:0 luma up to the 14- vs 20-bit weights; on pixel
    # pairs of one colour the chroma equals that colour's 4:2:0 chroma
    src = rng.integers(0, 256, (5, 16, 3), dtype=np.uint8)
    src[:, 1::2] = src[:, 0::2]
    err, buf = O.rgb_to_yuv(src, base)
    assert err == 0
    q = buf.reshape(5, 8, 4)
    yo, uo, vo = {"YUYV": (0, 1, 3), "UYVY": (1, 0, 2), "YVYU": (0, 3, 1)}[order]
    r, g, b = (src[:, 0::2, i].astype(np.int64) for i in range(3))
    u420 = (-155188 * r - 305135 * g + 460324 * b + (1 << 19) + (128 << 20)) >> 20
    v420 = (460324 * r - 385875 * g - 74448 * b + (1 << 19) + (128 << 20)) >> 20
    assert np.abs(q[..., uo].astype(int) - u420).max() <= 1
    assert np.abs(q[..., vo].astype(int) - v420).max() <= 1
    assert (q[..., yo] == q[..., yo + 2]).all()


def test_rgb_to_yuv_routing_errors():
    src = np.zeros((4, 10, 3), np.uint8)
    assert O.rgb_to_yuv(src[:, :9], 0x04)[0] == 1     # 4:2:2 odd width
    assert O.rgb_to_yuv(src, 0x94)[0] == 1            # VU and chroma-first
    assert O.rgb_to_yuv(src, 0x11)[0] == 1            # 4:4:4 with VU
    assert O.rgb_to_yuv(src, 0x02)[0] == 1            # NV12 through the one-buffer call
    assert O.rgb_to_yuv(src, 0x83)[0] == 1
    assert O.rgb_to_yuv_semiplanar(src, 0x03)[0] == 1


def test_row_stripes_equal_whole_calls(rng):
    # the stripe entry points the CPU baseline schedules: every band of
    # destination rows equals the same rows of the whole-image call
    L = O.lib()
    sz = C.c_size_t
    src = rng.integers(0, 256, (41, 57, 4), dtype=np.uint8)
    for dw, dh in ((114, 82), (28, 20), (43, 29)):
        _, whole = O.resize_linear(src, dw, dh)
        got = np.zeros_like(whole)
        for a, b in ((0, 5), (5, 6), (6, dh)):
            assert L.orc_resize_linear_u8_stripe(
                C.c_void_p(src.ctypes.data), sz(57 * 4), sz(57), sz(41), C.c_void_p(got.ctypes.data),
                sz(dw * 4), sz(dw), sz(dh), sz(4), sz(a), sz(b)) == 0
        np.testing.assert_array_equal(got, whole)
    t = np.array([[0.9, 0.1, 2.0], [-0.1, 0.95, 1.0], [1e-4, -1e-4, 1.0]], np.float32)
    _, whole = O.warp_perspective(src, t, 50, 30, "LINEAR", "REPLICATE")
    got = np.zeros_like(whole)
    bv = np.zeros(4, np.uint8)
    for a, b in ((0, 7), (7, 30)):
        assert L.orc_warp_perspective_u8_stripe(
            C.c_void_p(src.ctypes.data), sz(57 * 4), sz(57), sz(41), C.c_void_p(got.ctypes.data),
            sz(200), sz(50), sz(30), C.c_void_p(t.ctypes.data), sz(4), C.c_int(1), C.c_int(1),
            C.c_void_p(bv.ctypes.data), sz(a), sz(b)) == 0
    np.testing.assert_array_equal(got, whole)


# ---------------------------------------------------------------------------
# Exact pins of the NV12 chroma terms, the sigma tap table and the float
# gathers' rounding and contraction.
@pytest.mark.parametrize("fmt", [0x02, 0x22, 0x42, 0x62, 0x12, 0x32, 0x52, 0x72])
def test_nv12_equals_reference_test_formula(rng, fmt):
    # test_yuv420sp_to_rgb.cpp:112-160 with the weights of
    # yuv42x_coefficients.h:95-105, both read from the reference's files
    k = KATS["nv12_coefficients"]
    for h, w in [(2, 2), (7, 9), (16, 33), (31, 64)]:
        y = rng.integers(0, 256, (h, w), dtype=np.uint8)
        uv = rng.integers(0, 256, ((h + 1) // 2, 2 * ((w + 1) // 2)), dtype=np.uint8)
        err, out = O.yuv_sp_to_rgb(y, uv, fmt)
        assert err == 0
        np.testing.assert_array_equal(out, pins.nv12_reference_formula(y, uv, fmt, k))
    # every (U, V) pair once, with Y at both ends and in the middle: a wrong
    # weight or rounding constant changes some of these outputs
    uu, vv = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    uv = np.stack([uu, vv], axis=-1).reshape(256, 512).astype(np.uint8)
    for yl in (0, 16, 17, 128, 235, 255):
        y = np.full((512, 512), yl, np.uint8)
        err, out = O.yuv_sp_to_rgb(y, uv, fmt)
        np.testing.assert_array_equal(out, pins.nv12_reference_formula(y, uv, fmt, k))


def test_nv12_pin_rejects_a_mistyped_weight(rng):
    # the pin is discriminating: one weight off by one changes outputs
    k = dict(KATS["nv12_coefficients"])
    y = rng.integers(0, 256, (64, 64), dtype=np.uint8)
    uv = rng.integers(0, 256, (32, 64), dtype=np.uint8)
    _, out = O.yuv_sp_to_rgb(y, uv, 0x02)
    for name in ("rv", "gu", "gv", "bu"):
        bad = dict(k, **{name: k[name] + 1024})
        assert not np.array_equal(out, pins.nv12_reference_formula(y, uv, 0x02, bad)), name


def _half_kernel(k, sigma):
    import ctypes as C
    half = np.zeros(k // 2 + 1, np.uint8)
    ok = O.lib().orc_gaussian_half_kernel(half.ctypes.data, C.c_size_t(k // 2 + 1),
                                          C.c_float(sigma))
    return bool(ok), [int(v) for v in half]


# sigma values (size, sigma) where the float table and the double model
# round one entry differently, the value lying within 1e-4 of the half-way
# point; every other entry of every table on the grid is equal
SIGMA_BOUNDARY_CASES_LIMIT = 12


def test_sigma_half_kernel_equals_double_model_up_to_rounding_boundaries():
    grid = [0.0] + [round(0.05 * i, 2) for i in range(1, 241)]
    sizes = [3, 5, 7, 9, 15, 21, 11, 31, 63, 127, 235, 255]
    boundary = []
    for k in sizes:
        for sigma in grid:
            ok, table = _half_kernel(k, np.float32(sigma))
            want, near = pins.gaussian_half_kernel_double(k // 2 + 1, float(np.float32(sigma)),
                                                          follow=table)
            if not ok:
                assert want[-1] == 256, (k, sigma)  # all energy in the centre tap
                continue
            for i, (a, b) in enumerate(zip(table, want)):
                if a != b or (i < k // 2 and table[i] != want[i]):
                    assert near[i] < 1e-4, (k, sigma, i, a, b, near[i])
            # a followed entry is one whose rounding the two precisions split
            plain, _ = pins.gaussian_half_kernel_double(k // 2 + 1, float(np.float32(sigma)))
            if plain != table:
                boundary.append((k, sigma))
    assert len(boundary) <= SIGMA_BOUNDARY_CASES_LIMIT, boundary


def test_sigma_default_is_the_contracted_expression():
    # ksize * 0.15F + 0.35F rounds once on AArch64; at 235x235 the unfused
    # sigma gives different taps, so the oracle's table tells the two apart
    s_f, s_u = pins.default_sigma(235), pins.default_sigma(235, fused=False)
    assert s_f != s_u
    ok, table = _half_kernel(235, 0.0)
    assert ok
    assert table == _half_kernel(235, s_f)[1] != _half_kernel(235, s_u)[1]
    want, _ = pins.gaussian_half_kernel_double(118, s_f)
    assert table == want


def dyadic(rng, lo, hi, shape, bits=4):
    return (np.round(rng.uniform(lo, hi, shape) * (1 << bits)) / (1 << bits)).astype(np.float32)


@pytest.mark.parametrize("interp", ["LINEAR", "NEAREST"])
@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
@pytest.mark.parametrize("channels", [1, 2, 4])
def test_remap_f32_exact_on_dyadic_coordinates(rng, interp, border, channels):
    # with fractions of 1/16 every float step is exact, so fused and unfused
    # lerps agree and the result is the exact rational rounded half away from
    # zero: exact .5 ties (frequent here) tell FCVTA from ties-to-even and
    # truncation, negative coordinates tell floor from truncation
    h, w = 9, 13
    src = rng.integers(0, 256, (h, w, channels), dtype=np.uint8)
    mapx = dyadic(rng, -3, w + 2, (12, 20))
    mapy = dyadic(rng, -3, h + 2, (12, 20))
    bv = (11, 22, 33, 44)
    err, out = O.remap_f32(src, mapx, mapy, interp, border, bv)
    assert err == 0
    want = pins.remap_f32_exact(src, mapx, mapy, interp, border, bv)
    np.testing.assert_array_equal(out, want)
    # half-way fractions do occur, so the rounding mode is exercised
    assert (np.abs(mapx * 2 - np.round(mapx * 2)) == 0).mean() > 0.05


@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_remap_f32_special_coordinates(border):
    # NaN, infinities, ties, negative values and the far edge: results the
    # AArch64 conversion rules fix (NaN -> 0 for the taps and the lerp, ties
    # away from zero, saturating conversions)
    src = (np.arange(4 * 6).reshape(4, 6) * 9 + 3).astype(np.uint8)
    vals = [np.nan, np.inf, -np.inf, -0.5, -0.25, -1.0, 0.5, 1.5, 2.5, 4.75, 5.0, 5.5, 6.0,
            1e30, -1e30, 3.0]
    xs, ys = np.meshgrid(np.array(vals, np.float32), np.array(vals, np.float32))
    for interp in ("LINEAR", "NEAREST"):
        err, out = O.remap_f32(src, xs.astype(np.float32), ys.astype(np.float32), interp,
                               border, (77,))
        assert err == 0
        want = pins.remap_f32_exact(src, xs, ys, interp, border, (77,))
        np.testing.assert_array_equal(out, want)
    # spot values derived by hand: NaN gives 0 when interpolating
    _, out = O.remap_f32(src, np.full((1, 4), np.nan, np.float32), np.zeros((1, 4), np.float32),
                         "LINEAR", border, (77,))
    assert (out == 0).all()


def test_remap_f32_fuses_its_multiply_adds(rng):
    # The fused (vmla) and unfused lerps round differently in about one
    # pixel in 10^6, when a result lies within an ulp of a half-way point.
    # Such cases are found by a vectorised float32 search over 2x2 patches
    # (coordinates in [0, 1): full-precision fractions) and confirmed with
    # the exact model; on each, the oracle must give the fused result.
    srng = np.random.default_rng(5)
    n = 4_000_000
    a, b, c, d = (srng.integers(0, 256, n).astype(np.float32) for _ in range(4))
    x, y = srng.random(n, dtype=np.float32), srng.random(n, dtype=np.float32)

    def fused(p, q, r):
        return (p.astype(np.float64) * q.astype(np.float64) + r).astype(np.float32)

    def unfused(p, q, r):
        return (p * q).astype(np.float32) + r

    def lerp(fma):
        l0, l1 = fma(b - a, x, a), fma(d - c, x, c)
        return np.floor(fma((l1 - l0).astype(np.float32), y, l0).astype(np.float64) + 0.5)
    found = np.nonzero(lerp(fused) != lerp(unfused))[0]
    checked = 0
    for i in found:
        src = np.array([[a[i], b[i]], [c[i], d[i]]], np.uint8)
        mx = np.full((1, 4), x[i], np.float32)
        my = np.full((1, 4), y[i], np.float32)
        f = pins.remap_f32_exact(src, mx, my, "LINEAR", "REPLICATE", (0,), fused=True)
        u = pins.remap_f32_exact(src, mx, my, "LINEAR", "REPLICATE", (0,), fused=False)
        if np.array_equal(f, u):
            continue  # the float64 search rounded twice; not a real split
        _, out = O.remap_f32(src, mx, my, "LINEAR", "REPLICATE", (0,))
        np.testing.assert_array_equal(out, f)
        checked += 1
    assert checked >= 3, (len(found), checked)
    # and on ordinary coordinates the oracle equals the fused model
    src = rng.integers(0, 256, (16, 24), dtype=np.uint8)
    mapx = rng.uniform(-1, 24, (12, 20)).astype(np.float32)
    mapy = rng.uniform(-1, 16, (12, 20)).astype(np.float32)
    for border in ("REPLICATE", "CONSTANT"):
        _, out = O.remap_f32(src, mapx, mapy, "LINEAR", border, (5,))
        np.testing.assert_array_equal(
            out, pins.remap_f32_exact(src, mapx, mapy, "LINEAR", border, (5,)))


@pytest.mark.parametrize("interp", ["LINEAR", "NEAREST"])
@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_warp_exact_on_dyadic_transforms(rng, interp, border):
    # affine transforms with dyadic entries and w = 1 or 2: every step is
    # exact, so the result is fixed by the mathematics and the rounding rules
    src = rng.integers(0, 256, (20, 28, 3), dtype=np.uint8)
    for tw in (1.0, 2.0):
        T = np.array([[1.25 * tw, 0.375 * tw, -2.5 * tw], [-0.25 * tw, 0.875 * tw, 1.5 * tw],
                      [0.0, 0.0, tw]], np.float32)
        err, out = O.warp_perspective(src, T, 24, 18, interp, border, (9, 8, 7))
        assert err == 0
        np.testing.assert_array_equal(
            out, pins.warp_exact(src, T, 24, 18, interp, border, (9, 8, 7)))


@pytest.mark.parametrize("interp", ["LINEAR", "NEAREST"])
def test_warp_equals_fused_model(rng, interp):
    # perspective homographies: the oracle equals the exactly-rounded model
    # of the reference's operation sequence with fused multiply-adds
    src = rng.integers(0, 256, (30, 40), dtype=np.uint8)
    T = random_homography(rng, 40, 30, persp=2e-3)
    for border in ("REPLICATE", "CONSTANT"):
        _, out = O.warp_perspective(src, T, 36, 26, interp, border, (77,))
        np.testing.assert_array_equal(
            out, pins.warp_exact(src, T, 36, 26, interp, border, (77,)))


def test_threaded_oracle_helpers_equal_whole_calls(rng):
    # the full-size GPU tests check against the oracle's row stripes on a
    # thread pool; the stripes must give exactly the whole call
    src = rng.integers(0, 256, (133, 157, 3), dtype=np.uint8)
    kx, ky = [1, 2, 0, 1, 0, 0, 1], [0, 1, 0, 0, 1, 1, 0]
    T = random_homography(rng, 157, 133, persp=1e-3)
    pairs = [
        (O.gaussian_blur(src, 5, 0.0, "REFLECT"), O.gaussian_blur_mt(src, 5, 0.0, "REFLECT")),
        (O.median_blur(src, 5, "REVERSE"), O.median_blur_mt(src, 5, "REVERSE")),
        (O.separable_filter_2d(src, kx, ky), O.separable_filter_2d_mt(src, kx, ky)),
        (O.resize_linear(src, 301, 250), O.resize_linear_mt(src, 301, 250)),
        (O.warp_perspective(src, T, 140, 120, "LINEAR", "CONSTANT", (1, 2, 3)),
         O.warp_perspective_mt(src, T, 140, 120, "LINEAR", "CONSTANT", (1, 2, 3))),
    ]
    mxy = rng.integers(-10, 170, (120, 140, 2)).astype(np.int16)
    mfr = rng.integers(0, 1024, (120, 140)).astype(np.uint16)
    s1 = np.ascontiguousarray(src[..., 0])
    pairs.append((O.remap_s16point5(s1, mxy, mfr, "CONSTANT", (5,)),
                  O.remap_s16point5_mt(s1, mxy, mfr, "CONSTANT", (5,))))
    mx = rng.uniform(-5, 165, (120, 140)).astype(np.float32)
    my = rng.uniform(-5, 140, (120, 140)).astype(np.float32)
    pairs.append((O.remap_f32(src, mx, my), O.remap_f32_mt(src, mx, my)))
    y = rng.integers(0, 256, (133, 158), dtype=np.uint8)
    uv = rng.integers(0, 256, (67, 158), dtype=np.uint8)
    pairs.append((O.yuv_sp_to_rgb(y, uv, 0x52), O.yuv_sp_to_rgb_mt(y, uv, 0x52)))
    for i, ((e1, a), (e2, b)) in enumerate(pairs):
        assert e1 == e2 == 0, i
        np.testing.assert_array_equal(a, b, err_msg=str(i))
