"""Serving loop: continuous batching of image requests on one GPU.

Requests (an op on a frame that is already in device memory) arrive at any
time.  The scheduling runs in the library's native serving engine
(kleidicv_hip_server_*, csrc/serve.cpp); this module builds the requests'
op descriptors and hands the results back as torch views.

The engine keeps every buffer resident on the device and admits a request
once its output fits the HBM budget, taking the output from a device ring.
A large request (at least `large_px` output pixels) fills the GPU on its
own, so it is launched alone through its single-image device entry point on
the large-request slot, in arrival order; when more large requests back up
than the slot queues (the device is busy and the host sets the pace), up to
`large_batch` of them share one batch launch.  Small requests are packed
into batches (kleidicv_hip_batch_launch: one persistent launch per kernel
variant) on the other stream slots, so they never wait behind a large one
and the launch cost is shared (a lone small request runs through its own
entry point); the large-request slot also takes small batches when no large
request waits.  Each slot keeps up to `depth` launches queued on its stream,
so the device never waits for the host between them, and batches on
different slots overlap on the device.  A request completes when the host
observes its launch's event; its budget is then released.

Ops served: gaussian blur 5x5 (sigma 0, REVERSE), median blur 5x5
(REPLICATE), linear resize to ceil(w/2) x ceil(h/2), warp perspective
(LINEAR, REPLICATE, per-request homography) and NV12 to RGB.  Every op runs
the same kernels as its single-image entry point, so a request's result is
identical alone and in any mix.

The engine's host stand-in (standin_polls > 0) runs the same scheduler
without a GPU, for tests.
"""

import ctypes as C

import numpy as np

import kleidicv_hip as K

OPS = ("gaussian5", "median5", "resize_half", "warp", "nv12")


class Request:
    """One request.  `src` is an (h, w, c) device tensor; for nv12 it is the
    Y plane (h, w) and `uv` the interleaved chroma plane (h/2, w)."""

    __slots__ = ("rid", "op", "src", "uv", "T", "arrival", "dst", "admitted", "launched",
                 "done", "batch", "slot", "alone", "error", "sid", "_shape")

    def __init__(self, rid, op, src, uv=None, T=None, arrival=0.0):
        if op not in OPS:
            raise ValueError("unknown op %r" % op)
        self.rid, self.op, self.src, self.uv, self.T = rid, op, src, uv, T
        self.arrival = arrival
        self.dst = None
        self.admitted = self.launched = self.done = None
        self.batch = self.slot = self.alone = self.error = self.sid = None
        h, w = src.shape[0], src.shape[1]
        c = src.shape[2] if len(src.shape) == 3 else 1
        if op == "resize_half":
            self._shape = ((h + 1) // 2, (w + 1) // 2, c)
        elif op == "nv12":
            self._shape = (h, w, 3)
        else:
            self._shape = (h, w, c)

    def out_shape(self):
        return self._shape

    def out_bytes(self):
        s = self._shape
        return s[0] * s[1] * s[2]

    def out_pixels(self):
        s = self._shape
        return s[0] * s[1]

    def desc(self):
        """Op descriptor; without a dst the server takes the output from its ring."""
        dst = self.dst if self.dst is not None else self._shape
        if self.op == "gaussian5":
            return K.op_gaussian_blur(self.src, dst, 5, 0.0, "REVERSE")
        if self.op == "median5":
            return K.op_median_blur(self.src, dst, 5, "REPLICATE")
        if self.op == "resize_half":
            return K.op_resize_linear(self.src, dst)
        if self.op == "warp":
            return K.op_warp_perspective(self.src, dst, self.T, "LINEAR", "REPLICATE")
        return K.op_yuv_sp_to_rgb(self.src, self.uv, dst, K.NV12_TO_RGB)

    def run_alone(self, dst, stream=None):
        """The same op through its single-image device entry point."""
        if self.op == "gaussian5":
            K.gaussian_blur(self.src, 5, 0.0, "REVERSE", dst=dst, stream=stream, device_api=True)
        elif self.op == "median5":
            K.median_blur(self.src, 5, "REPLICATE", dst=dst, stream=stream, device_api=True)
        elif self.op == "resize_half":
            K.resize_linear(self.src, dst.shape[1], dst.shape[0], dst=dst, stream=stream,
                            device_api=True)
        elif self.op == "warp":
            K.warp_perspective(self.src, self.T, dst.shape[1], dst.shape[0], "LINEAR",
                               "REPLICATE", dst=dst, stream=stream, device_api=True)
        else:
            K.yuv_sp_to_rgb(self.src, self.uv, K.NV12_TO_RGB, dst=dst, stream=stream,
                            device_api=True)
        return dst


# kleidicv_hip_server_config / _request_info / _counters (kleidicv.h, section 5)
class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("workspace_reserve", C.c_size_t),
                ("large_px", C.c_size_t), ("slots", C.c_uint), ("max_batch_ops", C.c_uint),
                ("depth", C.c_uint), ("large_batch", C.c_uint), ("release_done", C.c_uint),
                ("plain", C.c_uint), ("standin_polls", C.c_uint), ("tuning", K.Tuning)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_uint64), ("state", C.c_int), ("error", C.c_int),
                ("slot", C.c_uint), ("alone", C.c_uint), ("launch", C.c_uint64),
                ("arrival", C.c_double), ("admitted", C.c_double), ("launched", C.c_double),
                ("done", C.c_double), ("dst", C.c_void_p), ("dst_stride", C.c_size_t),
                ("out_bytes", C.c_size_t)]


class Counters(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("pending", C.c_size_t),
                ("out_in_use", C.c_size_t), ("workspace_bytes", C.c_size_t),
                ("peak", C.c_size_t), ("stalled", C.c_uint)]


QUEUED, ADMITTED, LAUNCHED, DONE, FAILED = range(5)


def _native():
    L = K.lib()
    if not getattr(L, "_server_types", False):
        L.kleidicv_hip_server_clock.restype = C.c_double
        L.kleidicv_hip_server_pump.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_size_t)]
        L.kleidicv_hip_server_take_completed.restype = C.c_size_t
        L.kleidicv_hip_server_take_completed.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.kleidicv_hip_server_info.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.kleidicv_hip_server_release.argtypes = [C.c_void_p, C.c_uint64]
        L.kleidicv_hip_server_destroy.argtypes = [C.c_void_p]
        L._server_types = True
    return L


def clock():
    """The native server's monotonic clock, in seconds."""
    return _native().kleidicv_hip_server_clock()


class Server:
    """Continuous-batching server (the library's native engine,
    kleidicv_hip_server_*) over `slots` streams of one GPU.

    hbm_budget bounds the bytes of outputs and workspaces held at once (the
    inputs are resident already); each slot's work-list workspace is a fixed
    reserve of workspace_reserve bytes inside it.  Outputs come from a device
    ring of ring_bytes (default: what the budget leaves, at most 8 GiB).
    max_batch_ops bounds one batch; slot 0 is the large-request slot; depth
    bounds the launches queued on one slot's stream.  plain=True turns the
    server into the per-request baseline (every request alone on one stream
    at admission).  standin_polls > 0 runs the same scheduler against a host
    stand-in (no GPU): a launch completes on the n-th poll of its event.

    `clock` gives the times passed to the engine (default: its own
    monotonic clock).  Request fields admitted/launched/done/batch/slot/alone
    are filled in as requests complete; completed requests are appended to
    `completed`.
    """

    def __init__(self, torch=None, device=None, hbm_budget=64 << 30, slots=2,
                 max_batch_ops=512, workspace_reserve=64 << 20, tuning=None, large_px=1 << 18,
                 depth=4, release_done=False, large_batch=1, ring_bytes=None, plain=False,
                 standin_polls=0, clock=None):
        L = _native()
        self._L = L
        self.standin = standin_polls > 0
        self.clock = clock or L.kleidicv_hip_server_clock
        cfg = ServerConfig(hbm_budget=int(hbm_budget), workspace_reserve=int(workspace_reserve),
                           large_px=int(large_px), slots=int(slots),
                           max_batch_ops=int(max_batch_ops), depth=int(depth),
                           large_batch=int(large_batch), release_done=int(bool(release_done)),
                           plain=int(bool(plain)), standin_polls=int(standin_polls))
        if tuning is not None:
            cfg.tuning = tuning
        else:
            cfg.tuning = K.Tuning(blocks_per_cu=8, resize_band_rows=0, xcd_order=1, use_graph=0)
        if ring_bytes is None:
            ring_bytes = min(max(int(hbm_budget) - int(slots) * int(workspace_reserve), 0),
                             8 << 30)
        self.ring = None
        ring_ptr = 0
        if not self.standin and ring_bytes:
            self.ring = torch.empty(int(ring_bytes), dtype=torch.uint8, device=device)
            ring_ptr = self.ring.data_ptr()
        self._h = C.c_void_p(0)
        K._check(L.kleidicv_hip_server_create(C.byref(cfg), C.c_void_p(ring_ptr),
                                              C.c_size_t(int(ring_bytes)), C.byref(self._h)),
                 "kleidicv_hip_server_create")
        self.reqs = {}
        self.completed = []
        self._infos = (RequestInfo * 256)()

    # -- requests ---------------------------------------------------------------
    def submit(self, req):
        first = C.c_uint64(0)
        arr = C.c_double(req.arrival)
        K._check(self._L.kleidicv_hip_server_submit(self._h, C.byref(req.desc()),
                                                   C.c_size_t(1), C.byref(arr),
                                                   C.byref(first)),
                 "kleidicv_hip_server_submit")
        req.sid = first.value
        self.reqs[req.sid] = req

    def _complete(self, req, info):
        req.admitted = info.admitted
        req.launched = info.launched
        req.done = info.done
        req.batch = info.launch
        req.slot = info.slot
        req.alone = bool(info.alone)
        req.error = info.error if info.state == FAILED else None
        if self.ring is not None and info.dst and info.state == DONE:
            off = info.dst - self.ring.data_ptr()
            req.dst = self.ring[off:off + req.out_bytes()].view(req.out_shape())

    def pump(self):
        """One scheduling tick; returns the number of requests still pending."""
        pending = C.c_size_t(0)
        K._check(self._L.kleidicv_hip_server_pump(self._h, self.clock(), C.byref(pending)),
                 "kleidicv_hip_server_pump")
        while True:
            n = self._L.kleidicv_hip_server_take_completed(self._h, self._infos, 256)
            for k in range(n):
                info = self._infos[k]
                req = self.reqs.pop(info.id)
                self._complete(req, info)
                self.completed.append(req)
            if n < 256:
                break
        return pending.value

    def release(self, req):
        """Returns a completed request's output to the ring (release_done=False)."""
        K._check(self._L.kleidicv_hip_server_release(self._h, C.c_uint64(req.sid)),
                 "kleidicv_hip_server_release")

    def drain(self):
        """Pumps until every request has completed.  Raises when the rest
        waits for ring space that only release() can return."""
        while self.pump():
            if self.counters().stalled:
                raise RuntimeError("requests wait for ring space held by completed "
                                   "requests: release() their outputs")

    def replay(self, requests):
        """Real-time replay of `requests` (arrival offsets in seconds), all
        in the engine: returns the clock at the start."""
        order = sorted(requests, key=lambda r: r.arrival)
        n = len(order)
        descs = (K.OpDesc * n)(*[r.desc() for r in order])
        offs = (C.c_double * n)(*[r.arrival for r in order])
        res = (RequestInfo * n)()
        t0 = C.c_double(0)
        K._check(self._L.kleidicv_hip_server_replay(self._h, descs, C.c_size_t(n), offs, res,
                                                   C.byref(t0)),
                 "kleidicv_hip_server_replay")
        for r, info in zip(order, res):
            self._complete(r, info)
            self.completed.append(r)
        return t0.value

    # -- accounting ---------------------------------------------------------------
    def counters(self):
        c = Counters()
        K._check(self._L.kleidicv_hip_server_stats(self._h, C.byref(c)),
                 "kleidicv_hip_server_stats")
        return c

    @property
    def batches(self):
        return self.counters().launches

    @property
    def peak(self):
        return self.counters().peak

    @property
    def out_in_use(self):
        return self.counters().out_in_use

    @property
    def ws_bytes(self):
        return self.counters().workspace_bytes

    @property
    def in_use(self):
        c = self.counters()
        return c.out_in_use + c.workspace_bytes

    def pending(self):
        return self.counters().pending

    def close(self):
        if self._h:
            self._L.kleidicv_hip_server_destroy(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_arrivals(rng, n, qps):
    return np.cumsum(rng.exponential(1.0 / qps, n))


def run_served(server, requests):
    """Feeds `requests` (arrival offsets in seconds) to `server` in real time,
    inside the engine; returns the clock at the start."""
    return server.replay(requests)


def run_plain(torch, device, requests, hbm_budget=64 << 30, ring_bytes=None,
              release_done=False):
    """Baseline: each request launched alone through its single-image entry
    point at its arrival, on one stream, by the same native loop (outputs
    from the same kind of ring; a request whose output does not fit waits).
    Returns the clock at the start."""
    srv = Server(torch, device, hbm_budget=hbm_budget, slots=1, workspace_reserve=0,
                 plain=True, release_done=release_done, ring_bytes=ring_bytes)
    try:
        return srv.replay(requests)
    finally:
        srv.close()


def latency_stats(requests, t0):
    lat = np.array([(r.done - t0 - r.arrival) * 1e3 for r in requests])
    t_end = max(r.done for r in requests) - t0
    pix = sum(r.out_pixels() for r in requests)
    return {
        "requests": len(requests),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "max_ms": round(float(lat.max()), 3),
        "mpix_s": round(pix / t_end / 1e6, 1),
        "wall_s": round(t_end, 3),
    }
