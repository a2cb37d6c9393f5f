"""Seeded synthetic workloads for the batched and served paths.

Each builder returns plain Python records holding device tensors (torch, on
the given device) plus the per-image parameters, and can express the same
work three ways: as batch descriptors (one graph launch), as plain
per-image device-API calls, or as individual requests.

  pyramid_batch    uneven u8 images (W, H uniform in [lo, hi], C from
                   {1, 3, 4}, Gaussian kernel from {3, 5, 7}, border from the
                   four filter borders), each blurred and then reduced by a
                   three-level 0.5x linear-resize pyramid; level sizes are
                   ceil(previous / 2), as cv::pyrDown sizes them (with floor,
                   1-channel levels narrower than 32 bytes would need the 3:1
                   path, which the reference declines below 32 source bytes)
  transform_batch  RGBA frames, each warped through its own homography
                   (rotation <= 15 degrees about the centre, perspective terms
                   <= 1e-4) and, separately, remapped through its own float
                   maps (smooth displacement <= 20 px); bilinear, replicate
  serving_stream   requests with Poisson arrivals at a fixed rate, ops drawn
                   uniformly from the served set, frames from a resident
                   corpus whose sizes are log-uniform from 64x64 to
                   7680x4320 (one draw sets both axes) with 1, 3 or 4 channels
"""

import ctypes as C

import numpy as np

import kleidicv_hip as K

FILTER_BORDERS = ["REPLICATE", "REFLECT", "WRAP", "REVERSE"]


def _u8(torch, gen, dev, shape):
    return torch.randint(0, 256, shape, dtype=torch.uint8, device=dev, generator=gen)


def pyramid_batch(torch, dev, n=256, seed=3, lo=64, hi=2048, levels=3, padded=False):
    """Records {src, gauss, levels[], ksize, border} for n uneven images.

    With padded=True every other image lives in a wider allocation, so its
    rows are strided (ragged-stride coverage for tests).
    """
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    recs = []
    for i in range(n):
        w, h = int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))
        c = int(rng.choice([1, 3, 4]))
        k = int(rng.choice([3, 5, 7]))
        border = FILTER_BORDERS[int(rng.integers(0, 4))]
        if padded and i % 2 == 1:
            src = _u8(torch, gen, dev, (h, w + 3, c))[:, :w]
        else:
            src = _u8(torch, gen, dev, (h, w, c))
        gauss = torch.empty((h, w, c), dtype=torch.uint8, device=dev)
        lv, pw, ph = [], w, h
        for _ in range(levels):
            pw, ph = (pw + 1) // 2, (ph + 1) // 2
            lv.append(torch.empty((ph, pw, c), dtype=torch.uint8, device=dev))
        recs.append({"src": src, "gauss": gauss, "levels": lv, "ksize": k, "border": border})
    return recs


def pyramid_descs(recs):
    """Stage 0: Gaussian of every image; stage l: pyramid level l."""
    descs = []
    for r in recs:
        descs.append(K.op_gaussian_blur(r["src"], r["gauss"], r["ksize"], 0.0, r["border"],
                                        stage=0))
        prev = r["gauss"]
        for li, lv in enumerate(r["levels"]):
            descs.append(K.op_resize_linear(prev, lv, stage=li + 1))
            prev = lv
    return descs


def pyramid_plain(recs, stream):
    """The same work as per-image device-API launches, in image order."""
    for r in recs:
        K.gaussian_blur(r["src"], r["ksize"], 0.0, r["border"], dst=r["gauss"], stream=stream,
                        device_api=True)
        prev = r["gauss"]
        for lv in r["levels"]:
            K.resize_linear(prev, lv.shape[1], lv.shape[0], dst=lv, stream=stream,
                            device_api=True)
            prev = lv


def pyramid_counts(recs):
    """(destination pixels, algorithmic bytes) of one pass over the records."""
    pix = byts = 0
    for r in recs:
        h, w, c = r["src"].shape
        pix += h * w
        byts += 2 * h * w * c
        ph, pw = h, w
        for lv in r["levels"]:
            lh, lw = lv.shape[0], lv.shape[1]
            pix += lh * lw
            byts += (ph * pw + lh * lw) * c
            ph, pw = lh, lw
    return pix, byts


def homography(rng, w, h, max_deg=15.0, max_persp=1e-4):
    """Rotation about the image centre composed with a small projective term."""
    a = np.deg2rad(rng.uniform(-max_deg, max_deg))
    cx, cy = w / 2.0, h / 2.0
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Tc = np.array([[1, 0, cx], [0, 1, cy], [0, 0, 1]])
    Ti = np.array([[1, 0, -cx], [0, 1, -cy], [0, 0, 1]])
    P = np.array([[1, 0, 0], [0, 1, 0],
                  [rng.uniform(-max_persp, max_persp), rng.uniform(-max_persp, max_persp), 1]])
    T = (Tc @ R @ Ti) @ P
    return (T / T[2, 2]).astype(np.float32)


def plain_op_params(w, h, seed=7):
    """Seeded parameters of the plain 4K op set: 7x7 separable taps
    (integers in [0, 2]) and the warp homography (rotation <= 15 degrees,
    perspective terms <= 1e-4)."""
    rng = np.random.default_rng(seed)
    kx = rng.integers(0, 3, 7).astype(np.uint8)
    ky = rng.integers(0, 3, 7).astype(np.uint8)
    return kx, ky, homography(rng, w, h)


def smooth_maps(torch, gen, dev, w, h, amplitude):
    """Float maps x + dx, y + dy with |dx|, |dy| <= amplitude (sum of 4 waves)."""
    yy, xx = torch.meshgrid(torch.arange(h, device=dev, dtype=torch.float32),
                            torch.arange(w, device=dev, dtype=torch.float32), indexing="ij")
    r = torch.rand(16, generator=gen, device=dev, dtype=torch.float32)
    amp = amplitude / 4.0
    dx = sum(amp * torch.sin((0.002 + 0.02 * r[4 * i]) * xx + (0.002 + 0.02 * r[4 * i + 1]) * yy
                             + 6.28 * r[4 * i + 2]) for i in range(4))
    dy = sum(amp * torch.cos((0.002 + 0.02 * r[4 * i + 1]) * xx + (0.002 + 0.02 * r[4 * i]) * yy
                             + 6.28 * r[4 * i + 3]) for i in range(4))
    return (xx + dx).contiguous(), (yy + dy).contiguous()


def s16point5_maps(torch, gen, dev, w, h, amplitude=40.0):
    """OpenCV convertMaps-style fixed-point maps of a smooth displacement
    (sum of 4 seeded waves, |dx|, |dy| <= amplitude): interleaved s16 (x, y)
    integer parts and u16 5-bit fractions (fy << 5 | fx)."""
    yy, xx = torch.meshgrid(torch.arange(h, device=dev, dtype=torch.float64),
                            torch.arange(w, device=dev, dtype=torch.float64), indexing="ij")
    r = torch.rand(16, generator=gen, device=dev, dtype=torch.float64)
    amp = amplitude / 4.0
    fx = sum(amp * torch.sin((0.001 + 0.01 * r[4 * i]) * xx + (0.001 + 0.01 * r[4 * i + 1]) * yy
                             + 6.28 * r[4 * i + 2]) for i in range(4))
    fy = sum(amp * torch.cos((0.001 + 0.01 * r[4 * i + 1]) * xx + (0.001 + 0.01 * r[4 * i]) * yy
                             + 6.28 * r[4 * i + 3]) for i in range(4))
    ix = torch.floor((xx + fx) * 32).to(torch.int64)
    iy = torch.floor((yy + fy) * 32).to(torch.int64)
    mapxy = torch.stack([ix >> 5, iy >> 5], dim=-1).to(torch.int16).contiguous()
    mapfrac = ((ix & 31) | ((iy & 31) << 5)).to(torch.int16).contiguous()
    return mapxy, mapfrac


def transform_batch(torch, dev, n=512, seed=5, w=1280, h=720, channels=4, amplitude=20.0):
    """Records {warp_src, warp_dst, T, remap_src, remap_dst, mapx, mapy}."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    recs = []
    for _ in range(n):
        rec = {
            "warp_src": _u8(torch, gen, dev, (h, w, channels)),
            "warp_dst": torch.empty((h, w, channels), dtype=torch.uint8, device=dev),
            "T": homography(rng, w, h),
            "remap_src": _u8(torch, gen, dev, (h, w, channels)),
            "remap_dst": torch.empty((h, w, channels), dtype=torch.uint8, device=dev),
        }
        rec["mapx"], rec["mapy"] = smooth_maps(torch, gen, dev, w, h, amplitude)
        recs.append(rec)
    return recs


def transform_descs(recs):
    descs = []
    for r in recs:
        descs.append(K.op_warp_perspective(r["warp_src"], r["warp_dst"], r["T"], "LINEAR",
                                           "REPLICATE"))
        descs.append(K.op_remap_f32(r["remap_src"], r["remap_dst"], r["mapx"], r["mapy"],
                                    "LINEAR", "REPLICATE"))
    return descs


def transform_plain(recs, stream):
    for r in recs:
        h, w = r["warp_dst"].shape[:2]
        K.warp_perspective(r["warp_src"], r["T"], w, h, "LINEAR", "REPLICATE",
                           dst=r["warp_dst"], stream=stream, device_api=True)
        K.remap_f32(r["remap_src"], r["mapx"], r["mapy"], "LINEAR", "REPLICATE",
                    dst=r["remap_dst"], stream=stream, device_api=True)


def transform_counts(recs):
    """(destination pixels, algorithmic bytes): warp 2*4 B/px, remap 4+8+4 B/px."""
    pix = byts = 0
    for r in recs:
        h, w, c = r["warp_dst"].shape
        pix += 2 * h * w
        byts += 2 * h * w * c + (2 * h * w * c + 8 * h * w)
    return pix, byts


# ---------------------------------------------------------------------------
# Plain per-image launches as prepared C calls: the argument tuples are built
# once, so timing the loop measures the library's per-call cost and the
# kernels rather than Python wrapper overhead.
def _s(*v):
    return [C.c_size_t(int(x)) for x in v]


def _p(t):
    return C.c_void_p(t.data_ptr())


def _st(t):
    return t.stride(0) * t.element_size()


def pyramid_plain_calls(recs, stream):
    L = K.lib()
    sp = C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0))
    calls = []
    for r in recs:
        h, w, c = r["src"].shape
        calls.append((L.kleidicv_hip_gaussian_blur_u8,
                      (_p(r["src"]), *_s(_st(r["src"])), _p(r["gauss"]),
                       *_s(_st(r["gauss"]), w, h, c, r["ksize"], r["ksize"]), C.c_float(0.0),
                       C.c_float(0.0), C.c_int(K.BORDER[r["border"]]), sp)))
        prev = r["gauss"]
        for lv in r["levels"]:
            ph, pw = prev.shape[:2]
            lh, lw = lv.shape[:2]
            calls.append((L.kleidicv_hip_resize_linear_u8,
                          (_p(prev), *_s(_st(prev), pw, ph), _p(lv), *_s(_st(lv), lw, lh, c),
                           sp)))
            prev = lv
    return calls


def transform_plain_calls(recs, stream):
    L = K.lib()
    sp = C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0))
    bv = (C.c_uint8 * 4)(0, 0, 0, 0)
    calls = []
    for r in recs:
        h, w, c = r["warp_dst"].shape
        t = (C.c_float * 9)(*[float(v) for v in np.asarray(r["T"]).reshape(9)])
        r["_T_c"] = t  # keep alive with the record
        calls.append((L.kleidicv_hip_warp_perspective_u8,
                      (_p(r["warp_src"]), *_s(_st(r["warp_src"]), w, h), _p(r["warp_dst"]),
                       *_s(_st(r["warp_dst"]), w, h), t, *_s(c), C.c_int(1), C.c_int(1), bv,
                       sp)))
        calls.append((L.kleidicv_hip_remap_f32_u8,
                      (_p(r["remap_src"]), *_s(_st(r["remap_src"]), w, h), _p(r["remap_dst"]),
                       *_s(_st(r["remap_dst"]), w, h, c), _p(r["mapx"]), *_s(_st(r["mapx"])),
                       _p(r["mapy"]), *_s(_st(r["mapy"])), C.c_int(1), C.c_int(1), bv, sp)))
    return calls


def run_calls(calls):
    bad = [i for i, (fn, args) in enumerate(calls) if fn(*args) != K.OK]
    if bad:
        raise K.KleidiCVError(K.NOT_IMPLEMENTED, "plain call %d" % bad[0])


# ---------------------------------------------------------------------------
# Serving stream
def frame_corpus(torch, dev, n=48, seed=9, lo=(64, 64), hi=(7680, 4320)):
    """Resident frames: even W, H log-uniform between lo and hi, C from
    {1, 3, 4}; each frame also carries an NV12 chroma plane of its size."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    corpus = []
    for _ in range(n):
        u = rng.uniform()
        w = int(round(lo[0] * (hi[0] / lo[0]) ** u)) // 2 * 2
        h = int(round(lo[1] * (hi[1] / lo[1]) ** u)) // 2 * 2
        c = int(rng.choice([1, 3, 4]))
        corpus.append({"src": _u8(torch, gen, dev, (h, w, c)),
                       "y": _u8(torch, gen, dev, (h, w)),
                       "uv": _u8(torch, gen, dev, (h // 2, w)),
                       "T": homography(rng, w, h)})
    return corpus


def serving_requests(corpus, n, qps, seed=10):
    """n requests with Poisson arrivals at `qps` over the corpus."""
    import serving as S
    rng = np.random.default_rng(seed)
    t = S.poisson_arrivals(rng, n, qps)
    reqs = []
    for i in range(n):
        f = corpus[int(rng.integers(0, len(corpus)))]
        op = S.OPS[int(rng.integers(0, len(S.OPS)))]
        if op == "nv12":
            reqs.append(S.Request(i, op, f["y"], uv=f["uv"], arrival=float(t[i])))
        else:
            reqs.append(S.Request(i, op, f["src"], T=f["T"], arrival=float(t[i])))
    return reqs
