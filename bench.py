"""Benchmark of the u8 image hot path on one MI355X.

A step is one pass of the full plain op set on 3840x2160 u8 images, one
image per launch, inputs resident in HBM:
  separable_filter_2d 7x7 (1 ch, seeded integer taps, REPLICATE)
  median_blur 5x5 (1 ch, REPLICATE)
  resize_linear 4 ch 2x up (-> 7680x4320) and 0.5x down (-> 1920x1080)
  remap_s16point5 (1 ch, smooth seeded displacement <= 40 px, CONSTANT)
  warp_perspective (4 ch, seeded homography: rotation <= 15 deg,
                    perspective terms <= 1e-4, LINEAR, REPLICATE)
  yuv_semiplanar_to_rgb (NV12 -> RGB)
Inputs rotate over several distinct sets (> 512 MB in total) so reads come
from HBM rather than the 256 MB Infinity Cache.  Kernel time per op is taken
inside the timed steps with HIP events on the stream the kernels run on.

Prints one JSON line (metric, value in destination Mpix/s, per-op GB/s and
% of the 8 TB/s HBM peak, the dominant kernel's roofline, a CPU baseline).
"""

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "yocto-mirrors_gitlab_kleidi_kleidicv_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
W, H = 3840, 2160
METRIC = "Mpix/s + % HBM peak per op; batched/served vs plain Mpix/s; p99 latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--sets", type=int, default=3, help="distinct input sets rotated")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ops", default="", help="comma-separated subset of ops (profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 device copies of 512 MiB (HBM counter calibration)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched workloads (pyramid and transform batches)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving-loop workload")
    ap.add_argument("--no-host-api", action="store_true",
                    help="skip the drop-in (host buffer) entry point timings")
    ap.add_argument("--serving-requests", type=int, default=1500)
    return ap.parse_args()


def make_inputs(torch, gen, dev):
    """One seeded input set for every op, generated on the device."""
    def u8(*shape):
        return torch.randint(0, 256, shape, dtype=torch.uint8, device=dev, generator=gen)

    s = {}
    s["sep_src"] = u8(H, W)
    s["med_src"] = u8(H, W)
    s["up_src"] = u8(H, W, 4)
    s["down_src"] = u8(H, W, 4)
    s["remap_src"] = u8(H, W)
    s["warp_src"] = u8(H, W, 4)
    s["nv12_y"] = u8(H, W)
    s["nv12_uv"] = u8(H // 2, W)
    # smooth displacement, amplitude <= 40 px: sum of 4 seeded sinusoids
    yy, xx = torch.meshgrid(torch.arange(H, device=dev, dtype=torch.float64),
                            torch.arange(W, device=dev, dtype=torch.float64), indexing="ij")
    r = torch.rand(16, generator=gen, device=dev, dtype=torch.float64)
    fx = sum(10.0 * torch.sin((0.001 + 0.01 * r[4 * i]) * xx + (0.001 + 0.01 * r[4 * i + 1]) * yy
                              + 6.28 * r[4 * i + 2]) for i in range(4))
    fy = sum(10.0 * torch.cos((0.001 + 0.01 * r[4 * i + 1]) * xx + (0.001 + 0.01 * r[4 * i]) * yy
                              + 6.28 * r[4 * i + 3]) for i in range(4))
    ix = torch.floor((xx + fx) * 32).to(torch.int64)
    iy = torch.floor((yy + fy) * 32).to(torch.int64)
    s["mapxy"] = torch.stack([ix >> 5, iy >> 5], dim=-1).to(torch.int16).contiguous()
    s["mapfrac"] = ((ix & 31) | ((iy & 31) << 5)).to(torch.int16).contiguous()
    # inputs of the extended (non-step) ops, drawn last
    s["i420"] = u8((3 * H) // 2, W)
    s["yuyv"] = u8(H, 2 * W)
    s["yuv444"] = u8(H, 3 * W)
    s["rgb"] = u8(H, W, 3)
    return s


def seeded_params(np):
    rng = np.random.default_rng(7)
    kx = rng.integers(0, 3, 7).astype(np.uint8)
    ky = rng.integers(0, 3, 7).astype(np.uint8)
    a = np.deg2rad(rng.uniform(-15, 15))
    cx, cy = W / 2, H / 2
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Tc = np.array([[1, 0, cx], [0, 1, cy], [0, 0, 1]])
    Ti = np.array([[1, 0, -cx], [0, 1, -cy], [0, 0, 1]])
    P = np.array([[1, 0, 0], [0, 1, 0],
                  [rng.uniform(-1e-4, 1e-4), rng.uniform(-1e-4, 1e-4), 1]])
    T = (Tc @ R @ Ti) @ P
    T = T / T[2, 2]
    return kx, ky, T.astype(np.float32)


# destination pixels and algorithmic (compulsory) bytes per launch
OPS = [
    # name, dst pixels, bytes
    ("separable_7x7_1ch", W * H, 2 * W * H),
    ("median_5x5_1ch", W * H, 2 * W * H),
    ("resize_2x_up_4ch", 4 * W * H, 4 * W * H + 16 * W * H),
    ("resize_0.5x_4ch", W * H // 4, 4 * W * H + W * H),
    ("remap_s16point5_1ch", W * H, 8 * W * H),
    ("warp_perspective_4ch", W * H, 8 * W * H),
    ("nv12_to_rgb", W * H, (W * H * 9) // 2),
]


# Library kernels launched by each op (substrings of the kernel names).
OP_KERNELS = {
    "separable_7x7_1ch": ["separable_kernel"],
    "median_5x5_1ch": ["median_kernel"],
    "resize_2x_up_4ch": ["resize_up2x4_kernel"],
    "resize_0.5x_4ch": ["resize_quarter_kernel"],
    "remap_s16point5_1ch": ["remap_s16point5_kernel"],
    "warp_perspective_4ch": ["warp_kernel"],
    "nv12_to_rgb": ["nv12_kernel"],
}


# Entry points beyond the headline step, timed one op at a time on the same
# 3840x2160 inputs: name, dst pixels, algorithmic bytes, call.
def extended_ops(K, s, out, stream):
    def bd(src, name):
        return lambda: K.blur_and_downsample(src, "REVERSE", dst=out[name], stream=stream,
                                             device_api=True)
    def yuv(src, name, fmt):
        return lambda: K.yuv_to_rgb(src, W, H, fmt, dst=out[name], stream=stream,
                                    device_api=True)
    return [
        ("i420_to_rgb", W * H, (W * H * 3) // 2 + 3 * W * H, yuv(s["i420"], "i420_to_rgb", 0x03)),
        ("yuyv_to_rgb", W * H, 2 * W * H + 3 * W * H, yuv(s["yuyv"], "yuyv_to_rgb", 0x04)),
        ("yuv444_to_rgba", W * H, 3 * W * H + 4 * W * H,
         yuv(s["yuv444"], "yuv444_to_rgba", 0x41)),
        ("rgb_to_nv12", W * H, 3 * W * H + (W * H * 3) // 2,
         lambda: K.rgb_to_yuv_semiplanar(s["rgb"], 0x02, y=out["rgb_to_nv12"][0],
                                         uv=out["rgb_to_nv12"][1], stream=stream,
                                         device_api=True)),
        ("gaussian_31x31_1ch", W * H, 2 * W * H,
         lambda: K.gaussian_blur(s["sep_src"], 31, 0.0, "REPLICATE", dst=out["gaussian_31x31_1ch"],
                                 stream=stream, device_api=True)),
        ("median_9x9_1ch", W * H, 2 * W * H,
         lambda: K.median_blur(s["med_src"], 9, "REPLICATE", dst=out["median_9x9_1ch"],
                               stream=stream, device_api=True)),
        ("blur_and_downsample_1ch", (W // 2) * (H // 2), W * H + (W // 2) * (H // 2),
         bd(s["sep_src"], "blur_and_downsample_1ch")),
        ("blur_and_downsample_4ch", (W // 2) * (H // 2), 4 * (W * H + (W // 2) * (H // 2)),
         bd(s["down_src"], "blur_and_downsample_4ch")),
    ]


def extended_outputs(torch, dev):
    return {
        "i420_to_rgb": torch.empty(H, W, 3, dtype=torch.uint8, device=dev),
        "yuyv_to_rgb": torch.empty(H, W, 3, dtype=torch.uint8, device=dev),
        "yuv444_to_rgba": torch.empty(H, W, 4, dtype=torch.uint8, device=dev),
        "rgb_to_nv12": (torch.empty(H, W, dtype=torch.uint8, device=dev),
                        torch.empty(H // 2, W, dtype=torch.uint8, device=dev)),
        "gaussian_31x31_1ch": torch.empty(H, W, dtype=torch.uint8, device=dev),
        "median_9x9_1ch": torch.empty(H, W, dtype=torch.uint8, device=dev),
        "blur_and_downsample_1ch": torch.empty(H // 2, W // 2, dtype=torch.uint8, device=dev),
        "blur_and_downsample_4ch": torch.empty(H // 2, W // 2, 4, dtype=torch.uint8, device=dev),
    }


EXTENDED_INPUTS = ("i420", "yuyv", "yuv444", "rgb", "sep_src", "med_src", "down_src")


def bench_extended(K, torch, sets, dev, stream, steps, warmup):
    """Each extended op alone: `steps` launches between two events on the
    launch stream; ms per launch.  Every launch, warmup included, reads its
    own copy of the inputs (copies of the input sets as needed), and a 512 MiB
    fill before each op evicts the Infinity Cache (256 MB), so no timed launch
    finds its input cached by an earlier one: inputs come from HBM, as in
    the step."""
    out = extended_outputs(torch, dev)
    ext = [{k: s[k] for k in EXTENDED_INPUTS} for s in sets]
    while len(ext) < warmup + steps:
        ext.append({k: v.clone() for k, v in ext[len(ext) % len(sets)].items()})
    scrub = torch.empty(512 << 20, dtype=torch.uint8, device=dev)
    res = {}
    names = [o[0] for o in extended_ops(K, ext[0], out, stream)]
    for k, name in enumerate(names):
        calls = [extended_ops(K, s, out, stream)[k] for s in ext]
        _, pix, nbytes, _ = calls[0]
        scrub.fill_(k)
        torch.cuda.synchronize()
        for i in range(warmup):
            calls[i][3]()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for i in range(steps):
            calls[warmup + i][3]()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        gbs = nbytes / (ms * 1e-3) / 1e9
        res[name] = {"ms": round(ms, 4), "mpix_s": round(pix / (ms * 1e-3) / 1e6, 1),
                     "gb_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                     "bytes": nbytes}
    return res


def _pmc_pass(exe, counter, extra, td, tag):
    """One rocprofv3 --pmc run of this script; mean counter value per kernel name."""
    out = os.path.join(td, counter + "_" + tag)
    env = dict(os.environ, TMPDIR="/tmp")
    cmd = [exe, "--pmc", counter, "-f", "csv", "-d", out, "-o", "pmc", "--",
           sys.executable, os.path.abspath(__file__)] + extra
    subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, timeout=240, check=True)
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    acc = {}
    for f in files:
        for r in csv.DictReader(open(f)):
            acc.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def measure_traffic(args, ops):
    """HBM bytes per launch of each op from two rocprofv3 --pmc passes.

    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are read in separate runs of
    this script over the same ops, and each is divided by what the same
    counter reads per byte for a plain 512 MiB device copy.  Runs before this
    process touches the GPU, so the profiled children never share it with
    the timed run.
    """
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    names = ",".join(o[0] for o in ops)
    op_args = ["--steps", "2", "--warmup", "1", "--sets", str(args.sets), "--no-cpu-baseline",
               "--no-traffic", "--ops", names]
    try:
        with tempfile.TemporaryDirectory(dir="/tmp") as td:
            per = {}
            for counter in ("FETCH_SIZE", "WRITE_SIZE"):
                copy = _pmc_pass(exe, counter, ["--copy-calibration"], td, "copy")
                kib = [v for k, v in copy.items() if "copyBuffer" in k or "copy" in k.lower()]
                if not kib:
                    return None, "copy kernel not seen by rocprofv3"
                factor = max(kib) * 1024.0 / (512 << 20)
                vals = _pmc_pass(exe, counter, op_args, td, "ops")
                for name, _, _ in ops:
                    tot = 0.0
                    for sub in OP_KERNELS[name]:
                        tot += sum(v for k, v in vals.items() if sub in k and "kcv::" in k)
                    per.setdefault(name, {})[counter] = tot * 1024.0 / factor
                per.setdefault("_factor", {})[counter] = round(factor, 4)
    except (subprocess.SubprocessError, OSError) as e:
        return None, "rocprofv3 pass failed: %s" % type(e).__name__
    return per, None


def run_op(K, torch, name, s, out, kx, ky, T, stream):
    if name == "separable_7x7_1ch":
        K.separable_filter_2d(s["sep_src"], kx, ky, "REPLICATE", dst=out[name], stream=stream,
                              device_api=True)
    elif name == "median_5x5_1ch":
        K.median_blur(s["med_src"], 5, "REPLICATE", dst=out[name], stream=stream, device_api=True)
    elif name == "resize_2x_up_4ch":
        K.resize_linear(s["up_src"], 2 * W, 2 * H, dst=out[name], stream=stream, device_api=True)
    elif name == "resize_0.5x_4ch":
        K.resize_linear(s["down_src"], W // 2, H // 2, dst=out[name], stream=stream,
                        device_api=True)
    elif name == "remap_s16point5_1ch":
        K.remap_s16point5(s["remap_src"], s["mapxy"], s["mapfrac"], "CONSTANT", (0, 0, 0, 0),
                          dst=out[name], stream=stream, device_api=True)
    elif name == "warp_perspective_4ch":
        K.warp_perspective(s["warp_src"], T, W, H, "LINEAR", "REPLICATE", dst=out[name],
                           stream=stream, device_api=True)
    elif name == "nv12_to_rgb":
        K.yuv_sp_to_rgb(s["nv12_y"], s["nv12_uv"], K.NV12_TO_RGB, dst=out[name], stream=stream,
                        device_api=True)


def host_api(K, np, s_host, kx, ky, T, ops, reps=3):
    """The drop-in entry points on host (pageable numpy) buffers: each call
    copies its inputs to the device, runs, copies the result back and waits.
    Mpix/s including the host-device copies."""
    calls = {
        "separable_7x7_1ch": lambda: K.separable_filter_2d(s_host["sep_src"], kx, ky,
                                                           "REPLICATE"),
        "median_5x5_1ch": lambda: K.median_blur(s_host["med_src"], 5, "REPLICATE"),
        "resize_2x_up_4ch": lambda: K.resize_linear(s_host["up_src"], 2 * W, 2 * H),
        "resize_0.5x_4ch": lambda: K.resize_linear(s_host["down_src"], W // 2, H // 2),
        "remap_s16point5_1ch": lambda: K.remap_s16point5(
            s_host["remap_src"], s_host["mapxy"], s_host["mapfrac"].view(np.uint16), "CONSTANT"),
        "warp_perspective_4ch": lambda: K.warp_perspective(s_host["warp_src"], T, W, H, "LINEAR",
                                                           "REPLICATE"),
        "nv12_to_rgb": lambda: K.yuv_sp_to_rgb(s_host["nv12_y"], s_host["nv12_uv"], 0x02),
    }
    out, t_all, pix_all = {}, 0.0, 0
    for name, dst_pix, _ in ops:
        calls[name]()  # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            calls[name]()
        t = (time.perf_counter() - t0) / reps
        out[name] = {"ms": round(t * 1e3, 3), "mpix_s": round(dst_pix / t / 1e6, 1)}
        t_all += t
        pix_all += dst_pix
    out["all_ops_mpix_s"] = round(pix_all / t_all / 1e6, 1)
    return out


def cpu_baseline(np, s_host, kx, ky, T, ops, min_seconds=10.0, max_passes=4):
    """The C++ oracle on the host cores, each op split into row stripes run by
    a thread pool (the stripe scheduling of kleidicv_thread.cpp:60-75 with one
    stripe per thread): full 4K passes of the op set, repeated until about
    min_seconds of wall time (at most max_passes)."""
    import oracle_lib as O
    threads = max(1, min(16, os.cpu_count() or 1))
    calls = _cpu_calls(O, np, s_host, kx, ky, T)
    t_total, pix, passes = 0.0, 0, 0
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=threads) as pool:
        while passes < max_passes and (passes == 0 or t_total < min_seconds):
            passes += 1
            for name, dst_pix, _ in ops:
                rows, fn = calls[name]
                bands = [(rows * i // threads, rows * (i + 1) // threads) for i in range(threads)]
                t0 = time.perf_counter()
                errs = list(pool.map(lambda b: fn(*b), [b for b in bands if b[1] > b[0]]))
                t_total += time.perf_counter() - t0
                assert all(e == 0 for e in errs), (name, errs)
                pix += dst_pix
    return {"value": round(pix / t_total / 1e6, 2), "unit": "Mpix/s", "cores": threads,
            "kind": "port",
            "sample": "%d full 3840x2160 pass(es) of each of the %d ops, C++ oracle in %d "
                      "row stripes on %d threads (%.1f s)" % (passes, len(ops), threads,
                                                              threads, t_total)}


def _cpu_calls(O, np, s, kx, ky, T):
    """Per op: (rows to stripe, fn(y_begin, y_end) -> error) over the oracle's
    row-stripe entry points, or row-offset pointers where an op's output rows
    are independent (remap, NV12 row pairs)."""
    import ctypes as C
    L = O.lib()
    sz = C.c_size_t

    def P(a, off=0):
        return C.c_void_p(a.ctypes.data + off)
    kxa = np.ascontiguousarray(kx, np.uint8)
    kya = np.ascontiguousarray(ky, np.uint8)
    sep_src, med_src = s["sep_src"], s["med_src"]
    sep_dst, med_dst = np.empty_like(sep_src), np.empty_like(med_src)
    up_src, down_src = s["up_src"], s["down_src"]
    up_dst = np.empty((2 * H, 2 * W, 4), np.uint8)
    down_dst = np.empty((H // 2, W // 2, 4), np.uint8)
    rsrc, mxy, mfr = s["remap_src"], s["mapxy"], s["mapfrac"].view(np.uint16)
    rdst = np.empty((H, W), np.uint8)
    bv0 = np.zeros(4, np.uint8)
    wsrc = s["warp_src"]
    wdst = np.empty((H, W, 4), np.uint8)
    t32 = np.ascontiguousarray(T, np.float32)
    ny, nuv = s["nv12_y"], s["nv12_uv"]
    ndst = np.empty((H, W, 3), np.uint8)
    return {
        "separable_7x7_1ch": (H, lambda a, b: L.orc_separable_filter_2d_u8_stripe(
            P(sep_src), sz(W), P(sep_dst), sz(W), sz(W), sz(H), sz(1), P(kxa), sz(len(kxa)),
            P(kya), sz(len(kya)), C.c_int(O.BORDER["REPLICATE"]), sz(a), sz(b))),
        "median_5x5_1ch": (H, lambda a, b: L.orc_median_blur_u8_stripe(
            P(med_src), sz(W), P(med_dst), sz(W), sz(W), sz(H), sz(1), sz(5),
            C.c_int(O.BORDER["REPLICATE"]), sz(a), sz(b))),
        "resize_2x_up_4ch": (2 * H, lambda a, b: L.orc_resize_linear_u8_stripe(
            P(up_src), sz(4 * W), sz(W), sz(H), P(up_dst), sz(8 * W), sz(2 * W), sz(2 * H),
            sz(4), sz(a), sz(b))),
        "resize_0.5x_4ch": (H // 2, lambda a, b: L.orc_resize_linear_u8_stripe(
            P(down_src), sz(4 * W), sz(W), sz(H), P(down_dst), sz(2 * W), sz(W // 2),
            sz(H // 2), sz(4), sz(a), sz(b))),
        "remap_s16point5_1ch": (H, lambda a, b: L.orc_remap_s16point5_u8(
            P(rsrc), sz(W), sz(W), sz(H), P(rdst, a * W), sz(W), sz(W), sz(b - a), sz(1),
            P(mxy, a * 4 * W), sz(4 * W), P(mfr, a * 2 * W), sz(2 * W),
            C.c_int(O.BORDER["CONSTANT"]), P(bv0))),
        "warp_perspective_4ch": (H, lambda a, b: L.orc_warp_perspective_u8_stripe(
            P(wsrc), sz(4 * W), sz(W), sz(H), P(wdst), sz(4 * W), sz(W), sz(H), P(t32), sz(4),
            C.c_int(O.INTERP["LINEAR"]), C.c_int(O.BORDER["REPLICATE"]), P(bv0), sz(a),
            sz(b))),
        # row pairs: a stripe of pairs [a, b) is luma rows [2a, 2b)
        "nv12_to_rgb": (H // 2, lambda a, b: L.orc_yuv_semiplanar_to_rgb_u8(
            P(ny, 2 * a * W), sz(W), P(nuv, a * W), sz(W), P(ndst, 2 * a * 3 * W), sz(3 * W),
            sz(W), sz(2 * (b - a)), C.c_int(0x02))),
    }


def bench_batched(K, torch, WL, dev, stream, steps, warmup):
    """Batched workloads: one batch run per step (a few persistent launches,
    one per segment) vs the same work as per-image device-API calls, both
    timed with HIP events on `stream`."""
    results = {}
    specs = [
        ("pyramid_256_uneven", "256 u8 images, W,H uniform in [64,2048], C in {1,3,4}, "
         "Gaussian k in {3,5,7} (sigma 0), 4 filter borders, then a 3-level 0.5x linear "
         "pyramid (level size ceil(s/2)); 1024 ops in 4 stages",
         lambda: WL.pyramid_batch(torch, dev, n=256, seed=3), WL.pyramid_descs,
         WL.pyramid_plain_calls, WL.pyramid_counts,
         lambda r: [r["gauss"]] + r["levels"]),
        ("transform_512_rgba_720p", "512 RGBA 1280x720 frames warped through per-frame "
         "homographies + 512 frames remapped through per-frame f32 maps (<= 20 px), "
         "LINEAR, REPLICATE; 1024 ops in 1 stage",
         lambda: WL.transform_batch(torch, dev, n=512, seed=5), WL.transform_descs,
         WL.transform_plain_calls, WL.transform_counts,
         lambda r: [r["warp_dst"], r["remap_dst"]]),
    ]
    ev = lambda: torch.cuda.Event(enable_timing=True)
    for name, what, build, descs_fn, calls_fn, counts_fn, outs_fn in specs:
        recs = build()
        torch.cuda.synchronize()
        pix, nbytes = counts_fn(recs)
        calls = calls_fn(recs, stream)
        batch = K.Batch(descs_fn(recs), keep=recs)
        tuned, tune_ms = batch.tune(stream, iterations=3)
        # identical results: plain once, snapshot, batched once, compare on device
        WL.run_calls(calls)
        torch.cuda.synchronize()
        snap = [t.clone() for r in recs for t in outs_fn(r)]
        for r in recs:
            for t in outs_fn(r):
                t.fill_(0x5A)
        batch.run(stream)
        torch.cuda.synchronize()
        identical = all(torch.equal(a, b) for a, b in
                        zip(snap, [t for r in recs for t in outs_fn(r)]))
        del snap
        for _ in range(warmup):
            batch.run(stream)
        e0, e1 = ev(), ev()
        e0.record(stream)
        for _ in range(steps):
            batch.run(stream)
        e1.record(stream)
        torch.cuda.synchronize()
        b_ms = e0.elapsed_time(e1) / steps
        for _ in range(max(1, warmup)):
            WL.run_calls(calls)
        e0, e1 = ev(), ev()
        e0.record(stream)
        for _ in range(steps):
            WL.run_calls(calls)
        e1.record(stream)
        torch.cuda.synchronize()
        p_ms = e0.elapsed_time(e1) / steps
        gbs = nbytes / (b_ms * 1e-3) / 1e9
        results[name] = {
            "workload": what, "ops": len(calls), "work_items": batch.work_items,
            "tuning": tuned.as_dict(), "tuned_ms": round(tune_ms, 4),
            "batched_ms": round(b_ms, 4), "plain_ms": round(p_ms, 4),
            "batched_mpix_s": round(pix / (b_ms * 1e-3) / 1e6, 1),
            "plain_mpix_s": round(pix / (p_ms * 1e-3) / 1e6, 1),
            "speedup": round(p_ms / b_ms, 2),
            "bytes": nbytes, "batched_gb_s": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "identical_to_plain": bool(identical),
        }
        batch.close()
        del recs, calls, batch
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    return results


def bench_serving(torch, dev, n, qps_list=(2000.0, 20000.0, 1e9)):
    """Poisson request streams through the continuous-batching server and,
    with the same arrivals, as plain per-request launches on one stream."""
    import serving as S
    import workloads as WL
    corpus = WL.frame_corpus(torch, dev, n=48, seed=9)
    torch.cuda.synchronize()
    # Outputs of both paths come from one 8 GB device ring and are released
    # at completion (nothing consumes them), so neither path pays an
    # allocator call per request.
    arena = 8 << 30
    be = S.TorchBackend(torch, dev, arena_bytes=arena)
    budget = 64 << 30
    stream = torch.cuda.Stream()
    # warm-up of both paths (kernel loading, allocator caches)
    slots, max_ops, large_px = 2, 16, 1 << 18
    srv = S.Server(be, hbm_budget=budget, slots=slots, max_batch_ops=max_ops,
                   large_px=large_px, release_done=True)
    S.run_served(srv, WL.serving_requests(corpus, 100, 1e6, seed=1))
    S.run_plain(be, WL.serving_requests(corpus, 100, 1e6, seed=2), stream, release_done=True)
    torch.cuda.synchronize()
    out = {"workload": "%d requests, Poisson arrivals; ops uniform over gaussian 5x5, median "
                       "5x5, resize 0.5x, warp_perspective, NV12->RGB; frames from a resident "
                       "corpus of 48, W x H log-uniform 64x64..7680x4320, C in {1,3,4}; "
                       "latency = arrival to host-observed completion" % n,
           "hbm_budget_gb": budget / 2 ** 30, "output_ring_gb": arena / 2 ** 30,
           "slots": slots, "max_batch_ops": max_ops, "large_px": large_px}
    for qps in qps_list:
        reqs = WL.serving_requests(corpus, n, qps, seed=11)
        srv = S.Server(be, hbm_budget=budget, slots=slots, max_batch_ops=max_ops,
                       large_px=large_px, release_done=True)
        t0 = S.run_served(srv, reqs)
        torch.cuda.synchronize()
        served = S.latency_stats(reqs, t0)
        served["batches"] = srv.batches
        served["peak_resident_gb"] = round(srv.peak / 2 ** 30, 3)
        del reqs, srv
        plain_reqs = WL.serving_requests(corpus, n, qps, seed=11)
        t0 = S.run_plain(be, plain_reqs, stream, release_done=True)
        torch.cuda.synchronize()
        plain = S.latency_stats(plain_reqs, t0)
        del plain_reqs
        torch.cuda.empty_cache()
        key = "burst" if qps >= 1e8 else "qps_%d" % int(qps)
        out[key] = {
            "served": served, "plain": plain,
            "mpix_s_ratio": round(served["mpix_s"] / plain["mpix_s"], 2)}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every op is per image and shards by "
              "image/request with no exchange; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    ops = [o for o in OPS if not args.ops or o[0] in args.ops.split(",")]
    traffic, traffic_note = None, "not measured (--no-traffic)"
    if not args.no_traffic and not args.copy_calibration:
        traffic, traffic_note = measure_traffic(args, ops)

    import numpy as np
    import torch

    if args.copy_calibration:
        # A plain copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch
        # scale the counters read for the image kernels.
        n = 512 << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "launches": 3}))
        return

    import kleidicv_hip as K

    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(1234)
    kx, ky, T = seeded_params(np)
    sets = [make_inputs(torch, gen, dev) for _ in range(args.sets)]
    out = {
        "separable_7x7_1ch": torch.empty(H, W, dtype=torch.uint8, device=dev),
        "median_5x5_1ch": torch.empty(H, W, dtype=torch.uint8, device=dev),
        "resize_2x_up_4ch": torch.empty(2 * H, 2 * W, 4, dtype=torch.uint8, device=dev),
        "resize_0.5x_4ch": torch.empty(H // 2, W // 2, 4, dtype=torch.uint8, device=dev),
        "remap_s16point5_1ch": torch.empty(H, W, dtype=torch.uint8, device=dev),
        "warp_perspective_4ch": torch.empty(H, W, 4, dtype=torch.uint8, device=dev),
        "nv12_to_rgb": torch.empty(H, W, 3, dtype=torch.uint8, device=dev),
    }
    # A dedicated stream: the library launches on exactly this stream and the
    # timing events are recorded on it.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    for i in range(args.warmup):
        for name, _, _ in ops:
            run_op(K, torch, name, sets[i % len(sets)], out, kx, ky, T, stream)
    torch.cuda.synchronize()

    # Whole-step time: every op of a step back to back on one stream, with
    # nothing between the launches.
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for i in range(args.steps):
        s = sets[i % len(sets)]
        for name, _, _ in ops:
            run_op(K, torch, name, s, out, kx, ky, T, stream)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = start.elapsed_time(stop) / args.steps

    # Per-op duration: the same steps again with an event recorded on the
    # launch stream between consecutive ops.  Each recorded event adds a
    # fixed delay (the next op is dispatched only after the timestamp is
    # written), which the marked steps show as their excess over the plain
    # steps above; that excess is shared evenly among the ops, so the per-op
    # times add up to the plain step time.  What remains of an op's time is
    # its kernels' duration plus the gap between two launches, the figure
    # rocprofv3 --kernel-trace reports (less that gap).
    marks = [[torch.cuda.Event(enable_timing=True) for _ in range(len(ops) + 1)]
             for _ in range(args.steps)]
    torch.cuda.synchronize()
    for i in range(args.steps):
        s = sets[i % len(sets)]
        marks[i][0].record(stream)
        for k, (name, _, _) in enumerate(ops):
            run_op(K, torch, name, s, out, kx, ky, T, stream)
            marks[i][k + 1].record(stream)
    torch.cuda.synchronize()
    marked = {}
    for k, (name, _, _) in enumerate(ops):
        marked[name] = sum(m[k].elapsed_time(m[k + 1]) for m in marks) / args.steps
    event_delay = max(0.0, (sum(marked.values()) - step_ms) / len(ops))
    op_ms = {name: marked[name] - event_delay for name in marked}

    per_op = {}
    total_pix = 0
    for name, dst_pix, nbytes in ops:
        ms = op_ms[name]
        gbs = nbytes / (ms * 1e-3) / 1e9
        per_op[name] = {"ms": round(ms, 4), "mpix_s": round(dst_pix / (ms * 1e-3) / 1e6, 1),
                        "gb_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                        "bytes": nbytes}
        if traffic is not None:
            per_op[name]["traffic"] = int(traffic[name]["FETCH_SIZE"] +
                                          traffic[name]["WRITE_SIZE"])
        total_pix += dst_pix
    value = total_pix / (step_ms * 1e-3) / 1e6
    dom = max(per_op, key=lambda n: per_op[n]["ms"])
    d = per_op[dom]
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "wall_s": round(wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded uniform u8 pixels, seeded maps and homography)",
        "config": {"workload": "full plain op set, one 3840x2160 u8 image per launch "
                               "(%d ops per step, %d rotating input sets)" % (len(ops), args.sets),
                   "ops": [o[0] for o in ops]},
        "per_op": per_op,
        "per_op_note": "kernel time of each op inside the timed steps: event-bracketed time "
                       "less the per-event delay %.4f ms (marked minus plain step time, per op)"
                       % event_delay,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": d["gb_s"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["hbm_frac"],
                     "traffic": d.get("traffic")},
    }
    if traffic is None:
        result["roofline"]["traffic_note"] = traffic_note
    else:
        result["roofline"]["traffic_note"] = (
            "HBM bytes per launch: rocprofv3 FETCH_SIZE + WRITE_SIZE, each divided by its "
            "reading per byte of a plain 512 MiB copy (factors %s)" % json.dumps(traffic["_factor"]))
    if not args.ops:
        result["extended"] = bench_extended(K, torch, sets, dev, stream, args.steps,
                                            args.warmup)
        result["extended_note"] = ("entry points outside the headline step, each timed alone "
                                   "on the same 3840x2160 inputs (launch-to-launch time)")
    if not args.no_batched and not args.ops:
        import workloads as WL
        result["batched"] = bench_batched(K, torch, WL, dev, stream, args.steps, args.warmup)
    if not args.no_serving and not args.ops:
        result["serving"] = bench_serving(torch, dev, args.serving_requests)
    s_host = {k: v.cpu().numpy() for k, v in sets[0].items()}
    if not args.ops and not args.no_host_api:
        result["including_host_copies"] = host_api(K, np, s_host, kx, ky, T, ops)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(np, s_host, kx, ky, T, ops)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
