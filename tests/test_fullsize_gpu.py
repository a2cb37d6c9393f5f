"""Parity at the benchmark workloads' own sizes.

Every case runs one headline workload at full size through the C ABI and
compares it with the CPU oracle at tolerance 0 (the oracle's row-stripe
entry points on a thread pool, so a 4K or 8K comparison takes seconds), or,
for the batched and served paths, with the same work run image by image,
byte for byte, plus an oracle sample that includes the largest image.

  1920x1080 Gaussian 5x5 (sigma 0, REVERSE), host and device buffers
  3840x2160: separable 7x7, median 5x5, resize 4 ch 2x up (to 7680x4320)
             and 0.5x, remap s16point5 1 and 4 ch (smooth displacement of
             up to 40 px, CONSTANT), warp perspective 4 ch, NV12 to RGB
  256 uneven images (W, H in [64, 2048]) with a 3-level pyramid, batched
  512 + 512 RGBA 1280x720 warp and remap-f32 frames, batched
  served requests over frames from 64x64 to 7680x4320
  gathers whose source spans more than 2 GiB (64-bit address path)
"""

import numpy as np
import pytest

import kleidicv_hip as K
import oracle_lib as O
import serving as S
import workloads as WL

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

torch = pytest.importorskip("torch")

W4K, H4K = 3840, 2160


def dev():
    return torch.device("cuda:0")


def host(t):
    return t.cpu().numpy()


def assert_same(got, want, what=""):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    if not np.array_equal(got, want):
        bad = np.argwhere(got != want)
        i = tuple(bad[0])
        raise AssertionError("%s: %d mismatches, first at %s: got %d want %d" %
                             (what, len(bad), i, got[i], want[i]))


def rand_u8(gen, *shape, hi=256):
    return torch.randint(0, hi, shape, dtype=torch.uint8, device=dev(), generator=gen)


def generator(seed):
    g = torch.Generator(device=dev())
    g.manual_seed(seed)
    return g


# ---------------------------------------------------------------------------
# 1920x1080 Gaussian 5x5
def test_gaussian_5x5_1080p_reverse_host_and_device():
    rng = np.random.default_rng(20250601)
    src = rng.integers(0, 256, (1080, 1920), dtype=np.uint8)
    _, want = O.gaussian_blur_mt(src, 5, 0.0, "REVERSE")
    assert_same(K.gaussian_blur(src, 5, 0.0, "REVERSE"), want, "drop-in host buffers")
    d_src = torch.from_numpy(src).to(dev())
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    d_dst = K.gaussian_blur(d_src, 5, 0.0, "REVERSE", stream=stream, device_api=True)
    torch.cuda.synchronize()
    assert_same(host(d_dst), want, "device API")
    # drop-in entry point on device buffers written just before on torch's stream
    d_src2 = torch.from_numpy(src).to(dev())
    assert_same(host(K.gaussian_blur(d_src2, 5, 0.0, "REVERSE")), want, "drop-in device")


# ---------------------------------------------------------------------------
# 3840x2160 plain op set
def bench_taps():
    # the benchmark's seeded taps: integers in [0, 2] (no normalising shift
    # in the reference, so random u8 pixels saturate almost everywhere)
    kx, ky, _ = WL.plain_op_params(W4K, H4K)
    return kx, ky


@pytest.mark.parametrize("case", ["bench_taps", "unsaturated"])
def test_separable_7x7_4k(case):
    g = generator(101)
    if case == "bench_taps":
        kx, ky = bench_taps()
        src = rand_u8(g, H4K, W4K)
    else:
        # pixels in [0, 15] and taps summing to 4 each: every output is the
        # full sum, 16 * 15 = 240 at most, so no output saturates
        kx = np.array([1, 0, 1, 0, 1, 0, 1], np.uint8)
        ky = np.array([0, 2, 0, 1, 0, 1, 0], np.uint8)
        src = rand_u8(g, H4K, W4K, hi=16)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    got = K.separable_filter_2d(src, kx, ky, "REPLICATE", stream=stream, device_api=True)
    torch.cuda.synchronize()
    _, want = O.separable_filter_2d_mt(host(src), kx, ky, "REPLICATE")
    if case == "unsaturated":
        assert want.max() < 255
    assert_same(host(got), want, "separable 7x7 " + case)


@pytest.mark.parametrize("border", ["REPLICATE", "REVERSE"])
def test_median_5x5_4k(border):
    g = generator(102)
    src = rand_u8(g, H4K, W4K)
    torch.cuda.synchronize()
    got = K.median_blur(src, 5, border)  # drop-in entry point on device buffers
    _, want = O.median_blur_mt(host(src), 5, border)
    assert_same(host(got), want, "median 5x5 " + border)


def test_resize_2x_up_4ch_4k_to_8k():
    g = generator(103)
    src = rand_u8(g, H4K, W4K, 4)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    got = K.resize_linear(src, 2 * W4K, 2 * H4K, stream=stream, device_api=True)
    torch.cuda.synchronize()
    _, want = O.resize_linear_mt(host(src), 2 * W4K, 2 * H4K)
    assert_same(host(got), want, "resize 2x up 4ch")


def test_resize_half_4ch_4k():
    g = generator(104)
    src = rand_u8(g, H4K, W4K, 4)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    got = K.resize_linear(src, W4K // 2, H4K // 2, stream=stream, device_api=True)
    torch.cuda.synchronize()
    _, want = O.resize_linear_mt(host(src), W4K // 2, H4K // 2)
    assert_same(host(got), want, "resize 0.5x 4ch")


@pytest.mark.parametrize("channels", [1, 4])
@pytest.mark.parametrize("border", ["CONSTANT", "REPLICATE"])
def test_remap_s16point5_4k_displacement_40px(channels, border):
    g = generator(105 + channels)
    shape = (H4K, W4K) if channels == 1 else (H4K, W4K, channels)
    src = rand_u8(g, *shape)
    mapxy, mapfrac = WL.s16point5_maps(torch, g, dev(), W4K, H4K, amplitude=40.0)
    bv = (17, 99, 201, 3)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    got = K.remap_s16point5(src, mapxy, mapfrac, border, bv, stream=stream, device_api=True)
    torch.cuda.synchronize()
    mxy = host(mapxy)
    # the displacement reaches outside the frame, so the border is exercised
    assert mxy[..., 0].min() < 0 and mxy[..., 0].max() >= W4K
    _, want = O.remap_s16point5_mt(host(src), mxy, host(mapfrac).view(np.uint16), border, bv)
    assert_same(host(got), want, "remap s16point5 %dch %s" % (channels, border))


@pytest.mark.parametrize("border", ["REPLICATE", "CONSTANT"])
def test_warp_perspective_4ch_4k(border):
    g = generator(107)
    src = rand_u8(g, H4K, W4K, 4)
    _, _, T = WL.plain_op_params(W4K, H4K)  # the benchmark's homography
    bv = (5, 6, 7, 8)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    got = K.warp_perspective(src, T, W4K, H4K, "LINEAR", border, bv, stream=stream,
                             device_api=True)
    torch.cuda.synchronize()
    _, want = O.warp_perspective_mt(host(src), T, W4K, H4K, "LINEAR", border, bv)
    assert_same(host(got), want, "warp 4ch " + border)


def test_nv12_to_rgb_4k():
    g = generator(108)
    y, uv = rand_u8(g, H4K, W4K), rand_u8(g, H4K // 2, W4K)
    torch.cuda.synchronize()
    for fmt in (K.NV12_TO_RGB, K.NV21_TO_BGRA):
        got = K.yuv_sp_to_rgb(y, uv, fmt)  # drop-in data symbol on device buffers
        _, want = O.yuv_sp_to_rgb_mt(host(y), host(uv), fmt)
        assert_same(host(got), want, "nv12 fmt 0x%x" % fmt)


# ---------------------------------------------------------------------------
# Batched workloads at full size
def test_pyramid_batch_256_uneven_full_size():
    recs = WL.pyramid_batch(torch, dev(), n=256, seed=3)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    WL.pyramid_plain(recs, stream)
    torch.cuda.synchronize()
    want = [[r["gauss"].clone()] + [lv.clone() for lv in r["levels"]] for r in recs]
    for r in recs:
        r["gauss"].fill_(0xA5)
        for lv in r["levels"]:
            lv.fill_(0x5A)
    torch.cuda.synchronize()
    with K.Batch(WL.pyramid_descs(recs), keep=recs) as b:
        b.run(stream)
        torch.cuda.synchronize()
    for i, r in enumerate(recs):
        for k, (got, w) in enumerate(zip([r["gauss"]] + r["levels"], want[i])):
            if not torch.equal(got, w):
                assert_same(host(got), host(w), "image %d output %d" % (i, k))
    # the oracle on 8 images: the largest, the smallest and six more
    areas = [r["src"].numel() for r in recs]
    pick = sorted({int(np.argmax(areas)), int(np.argmin(areas)), 0, 1, 2, 100, 200, 255})
    for i in pick:
        r = recs[i]
        _, g = O.gaussian_blur_mt(host(r["src"]), r["ksize"], 0.0, r["border"])
        assert_same(host(r["gauss"]), g, "oracle gaussian image %d" % i)
        prev = g
        for k, lv in enumerate(r["levels"]):
            err, want_lv = O.resize_linear_mt(prev, lv.shape[1], lv.shape[0])
            assert err == 0
            assert_same(host(lv), want_lv, "oracle level %d image %d" % (k, i))
            prev = want_lv


def test_transform_batch_512_rgba_720p_full_size():
    recs = WL.transform_batch(torch, dev(), n=512, seed=5)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    WL.transform_plain(recs, stream)
    torch.cuda.synchronize()
    want = [(r["warp_dst"].clone(), r["remap_dst"].clone()) for r in recs]
    for r in recs:
        r["warp_dst"].fill_(0x11)
        r["remap_dst"].fill_(0x22)
    torch.cuda.synchronize()
    with K.Batch(WL.transform_descs(recs), keep=recs) as b:
        b.run(stream)
        torch.cuda.synchronize()
    for i, r in enumerate(recs):
        if not torch.equal(r["warp_dst"], want[i][0]):
            assert_same(host(r["warp_dst"]), host(want[i][0]), "warp %d" % i)
        if not torch.equal(r["remap_dst"], want[i][1]):
            assert_same(host(r["remap_dst"]), host(want[i][1]), "remap %d" % i)
    del want
    for i in (0, 255, 511):
        r = recs[i]
        _, ww = O.warp_perspective_mt(host(r["warp_src"]), r["T"], 1280, 720, "LINEAR",
                                      "REPLICATE")
        assert_same(host(r["warp_dst"]), ww, "oracle warp %d" % i)
        _, wr = O.remap_f32_mt(host(r["remap_src"]), host(r["mapx"]), host(r["mapy"]),
                               "LINEAR", "REPLICATE")
        assert_same(host(r["remap_dst"]), wr, "oracle remap %d" % i)


# ---------------------------------------------------------------------------
# Serving over the full frame-size range
def served_oracle(r):
    src = host(r.src)
    if r.op == "gaussian5":
        return O.gaussian_blur_mt(src, 5, 0.0, "REVERSE")[1]
    if r.op == "median5":
        return O.median_blur_mt(src, 5, "REPLICATE")[1]
    if r.op == "resize_half":
        h, w, _ = r.out_shape()
        return O.resize_linear_mt(src, w, h)[1]
    if r.op == "warp":
        h, w, _ = r.out_shape()
        return O.warp_perspective_mt(src, r.T, w, h, "LINEAR", "REPLICATE")[1]
    return O.yuv_sp_to_rgb_mt(src, host(r.uv), K.NV12_TO_RGB)[1]


def test_served_full_size_range_equals_alone():
    # log-uniform sizes from 64x64 to 7680x4320, plus both ends themselves
    corpus = (WL.frame_corpus(torch, dev(), n=22, seed=51) +
              WL.frame_corpus(torch, dev(), n=1, seed=53, lo=(7680, 4320), hi=(7680, 4320)) +
              WL.frame_corpus(torch, dev(), n=1, seed=54, lo=(64, 64), hi=(64, 64)))
    torch.cuda.synchronize()
    reqs = WL.serving_requests(corpus, 200, qps=50000.0, seed=52)
    assert max(r.src.shape[0] * r.src.shape[1] for r in reqs) == 7680 * 4320
    # every output stays in the ring until the comparison below (the budget
    # is released as requests complete)
    srv = S.Server(torch, dev(), hbm_budget=16 << 30, slots=2, max_batch_ops=16,
                   large_px=1 << 18, ring_bytes=12 << 30)
    S.run_served(srv, reqs)
    torch.cuda.synchronize()
    assert len(srv.completed) == len(reqs)
    assert srv.batches > 1
    stream = torch.cuda.Stream()
    for r in reqs:
        a = r.run_alone(torch.empty(r.out_shape(), dtype=torch.uint8, device=dev()), stream)
        torch.cuda.synchronize()
        if not torch.equal(r.dst.reshape(a.shape), a):
            assert_same(host(r.dst).reshape(a.shape), host(a), "request %d (%s)" % (r.rid, r.op))
    # the oracle on the largest request of each op
    by_op = {}
    for r in reqs:
        n = r.src.numel()
        if r.op not in by_op or n > by_op[r.op].src.numel():
            by_op[r.op] = r
    for op, r in sorted(by_op.items()):
        want = served_oracle(r)
        assert_same(host(r.dst).reshape(want.shape), want, "oracle %s %s" % (op, tuple(r.src.shape)))


# ---------------------------------------------------------------------------
# Sources spanning 2 GiB or more: the gathers leave 32-bit buffer offsets
# for 64-bit addresses.
@pytest.mark.parametrize("channels", [1, 4])
def test_gathers_with_sources_beyond_2gib(channels):
    rng = np.random.default_rng(61 + channels)
    w, h = 96, 4
    stride = 0x30000000  # 768 MiB between rows: the source spans 2.25 GiB
    buf = torch.empty(h * stride, dtype=torch.uint8, device=dev())
    compact = rng.integers(0, 256, (h, w * channels), dtype=np.uint8)
    for y in range(h):
        buf[y * stride:y * stride + w * channels].copy_(torch.from_numpy(compact[y]))
    src = buf.as_strided((h, w, channels), (stride, channels, 1))
    src_np = compact.reshape(h, w, channels)
    dw, dh = 64, 16
    mx = rng.uniform(-3, w + 2, (dh, dw)).astype(np.float32)
    my = rng.uniform(-1.5, h + 0.5, (dh, dw)).astype(np.float32)
    mx[0, :4] = [w - 1, w - 1.5, 0.25, 0.0]  # rows and columns at the far edge
    my[0, :4] = [h - 1, h - 1.25, h - 1, 0.0]
    d_mx, d_my = torch.from_numpy(mx).to(dev()), torch.from_numpy(my).to(dev())
    T = np.array([[1.4, 0.05, -2.0], [0.01, 0.2, -0.5], [1e-4, 2e-4, 1.0]], np.float32)
    torch.cuda.synchronize()
    for interp in ("LINEAR", "NEAREST"):
        for border in ("REPLICATE", "CONSTANT"):
            bv = (9, 8, 7, 6)
            got = K.remap_f32(src, d_mx, d_my, interp, border, bv)
            _, want = O.remap_f32(src_np, mx, my, interp, border, bv)
            assert_same(host(got).reshape(want.shape), want, "remap f32 %s %s" % (interp, border))
            got = K.warp_perspective(src, T, dw, dh, interp, border, bv)
            _, want = O.warp_perspective(src_np, T, dw, dh, interp, border, bv)
            assert_same(host(got).reshape(want.shape), want, "warp %s %s" % (interp, border))
