"""Serving loop: continuous batching of image requests on one GPU.

Requests (an op on a frame that is already in device memory) arrive at any
time.  The server keeps every buffer resident on the device and admits a
request once its output fits the HBM budget.  Admitted requests are split by
size.  A large request (at least `large_px` output pixels) fills the GPU on
its own, so it is launched alone through its single-image device entry point
on the large-request slot, in arrival order.  Small requests are packed into
batches (kleidicv_hip_batch_launch: one persistent launch per kernel variant)
on the other stream slots, so they never wait behind a large one and the
launch cost is shared; the large-request slot also takes small batches when
no large request waits.  Batches on different slots overlap on the device,
so requests that arrive during a batch start before it ends.  A request
completes when the host observes its launch's event; its output is then
handed to the caller and its budget released.

Ops served: gaussian blur 5x5 (sigma 0, REVERSE), median blur 5x5
(REPLICATE), linear resize to ceil(w/2) x ceil(h/2), warp perspective
(LINEAR, REPLICATE, per-request homography) and NV12 to RGB.  Every op runs
the same kernels as its single-image entry point, so a request's result is
identical alone and in any mix.

The device side is reached through a backend object (TorchBackend for the
GPU), so the scheduling logic also runs against a host-side stand-in.
"""

import collections
import time

import numpy as np

import kleidicv_hip as K

OPS = ("gaussian5", "median5", "resize_half", "warp", "nv12")


class Request:
    """One request.  `src` is an (h, w, c) device tensor; for nv12 it is the
    Y plane (h, w) and `uv` the interleaved chroma plane (h/2, w)."""

    __slots__ = ("rid", "op", "src", "uv", "T", "arrival", "dst", "admitted", "launched",
                 "done", "batch")

    def __init__(self, rid, op, src, uv=None, T=None, arrival=0.0):
        if op not in OPS:
            raise ValueError("unknown op %r" % op)
        self.rid, self.op, self.src, self.uv, self.T = rid, op, src, uv, T
        self.arrival = arrival
        self.dst = None
        self.admitted = self.launched = self.done = None
        self.batch = None

    def out_shape(self):
        h, w = self.src.shape[0], self.src.shape[1]
        c = self.src.shape[2] if len(self.src.shape) == 3 else 1
        if self.op == "resize_half":
            return ((h + 1) // 2, (w + 1) // 2, c)
        if self.op == "nv12":
            return (h, w, 3)
        return (h, w, c)

    def out_bytes(self):
        s = self.out_shape()
        return s[0] * s[1] * s[2]

    def out_pixels(self):
        s = self.out_shape()
        return s[0] * s[1]

    def desc(self):
        if self.op == "gaussian5":
            return K.op_gaussian_blur(self.src, self.dst, 5, 0.0, "REVERSE")
        if self.op == "median5":
            return K.op_median_blur(self.src, self.dst, 5, "REPLICATE")
        if self.op == "resize_half":
            return K.op_resize_linear(self.src, self.dst)
        if self.op == "warp":
            return K.op_warp_perspective(self.src, self.dst, self.T, "LINEAR", "REPLICATE")
        return K.op_yuv_sp_to_rgb(self.src, self.uv, self.dst, K.NV12_TO_RGB)

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


class TorchBackend:
    """Device side through torch tensors, streams and events."""

    def __init__(self, torch, device):
        self.torch, self.device = torch, device

    def stream(self):
        return self.torch.cuda.Stream(device=self.device)

    def alloc(self, shape):
        return self.torch.empty(shape, dtype=self.torch.uint8, device=self.device)

    def workspace(self, nbytes):
        return self.torch.empty(max(int(nbytes), 1), dtype=self.torch.uint8,
                                device=self.device)

    def launch(self, descs, workspace, stream, tuning, requests):
        """Launches, or returns the workspace bytes needed (nothing launched)."""
        return K.batch_launch(descs, workspace, stream, tuning)

    def launch_one(self, request, stream):
        """One request through its single-image device entry point."""
        request.run_alone(request.dst, stream)

    def record(self, stream):
        ev = self.torch.cuda.Event()
        ev.record(stream)
        return ev

    def query(self, event):
        return event.query()


class Server:
    """Continuous-batching server over `slots` streams.

    hbm_budget bounds the bytes of outputs and workspaces held at once (the
    inputs are resident already).  Each slot's work-list workspace comes out
    of a fixed reserve of workspace_reserve bytes; outputs are admitted
    against the rest.  max_batch_ops bounds one batch.  Slot 0 is the
    large-request slot; with one slot, everything runs on it.
    """

    def __init__(self, backend, hbm_budget=64 << 30, slots=2, max_batch_ops=512,
                 workspace_reserve=64 << 20, tuning=None, clock=time.perf_counter,
                 large_px=1 << 18):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        if slots * workspace_reserve >= hbm_budget:
            raise ValueError("the workspace reserve leaves no budget for outputs")
        self.be = backend
        self.budget = int(hbm_budget)
        self.reserve = int(workspace_reserve)
        self.out_budget = self.budget - slots * self.reserve
        self.out_in_use = 0
        self.peak = 0
        self.max_batch_ops = max_batch_ops
        self.tuning = tuning
        self.clock = clock
        self.streams = [backend.stream() for _ in range(slots)]
        self.inflight = [None] * slots  # (event, requests)
        self.ws = [None] * slots
        self.ws_bytes = [0] * slots
        self.large_px = large_px
        self.queue = collections.deque()   # arrived, waiting for budget
        self.ready = collections.deque()   # admitted small requests, waiting for a slot
        self.ready_large = collections.deque()  # admitted large requests
        self.completed = []
        self.batches = 0

    # -- admission ----------------------------------------------------------
    def submit(self, req):
        self.queue.append(req)

    def _admit(self, now):
        while self.queue:
            req = self.queue[0]
            need = req.out_bytes()
            if need > self.out_budget:
                raise ValueError("request %d needs %d bytes, more than the whole budget"
                                 % (req.rid, need))
            if self.out_in_use + need > self.out_budget:
                return  # FIFO: later requests wait behind the head
            self.queue.popleft()
            self.out_in_use += need
            self.peak = max(self.peak, self.in_use)
            req.dst = self.be.alloc(req.out_shape())
            req.admitted = now
            if req.out_pixels() >= self.large_px:
                self.ready_large.append(req)
            else:
                self.ready.append(req)

    # -- completion -----------------------------------------------------------
    def _retire(self, now):
        for i, slot in enumerate(self.inflight):
            if slot is None or not self.be.query(slot[0]):
                continue
            _, reqs = slot
            for r in reqs:
                r.done = now
                self.out_in_use -= r.out_bytes()
                self.completed.append(r)
            self.inflight[i] = None

    # -- batching -----------------------------------------------------------
    def _launch(self, now):
        for i in range(len(self.streams)):
            if self.inflight[i] is not None:
                continue
            if i == 0 and self.ready_large:
                req = self.ready_large.popleft()
                self.be.launch_one(req, self.streams[0])
                self._started(0, [req], now)
                continue
            if not self.ready:
                continue
            reqs = []
            while self.ready and len(reqs) < self.max_batch_ops:
                reqs.append(self.ready.popleft())
            # A slot's workspace lives in its reserve.  Launch into it; when
            # the work lists need more, grow it within the reserve or halve
            # the batch (the rest goes back to the front of the queue).
            descs = [r.desc() for r in reqs]
            while True:
                need = self.be.launch(descs, self.ws[i], self.streams[i], self.tuning, reqs)
                if need == 0:
                    break
                if need <= self.reserve:
                    grow = min(max(need, 2 * self.ws_bytes[i]), self.reserve)
                    self.ws[i] = self.be.workspace(grow)
                    self.ws_bytes[i] = grow
                    self.peak = max(self.peak, self.in_use)
                    continue
                if len(reqs) == 1:
                    raise ValueError("work lists of request %d exceed the workspace reserve"
                                     % reqs[0].rid)
                keep = len(reqs) // 2
                self.ready.extendleft(reversed(reqs[keep:]))
                reqs, descs = reqs[:keep], descs[:keep]
            self._started(i, reqs, now)

    def _started(self, i, reqs, now):
        ev = self.be.record(self.streams[i])
        self.batches += 1
        for r in reqs:
            r.launched = now
            r.batch = self.batches
        self.inflight[i] = (ev, reqs)

    def pump(self):
        """One scheduling tick; returns the number of requests still pending."""
        now = self.clock()
        self._retire(now)
        self._admit(now)
        self._launch(now)
        return self.pending()

    @property
    def in_use(self):
        """Device bytes held now: admitted outputs plus workspaces."""
        return self.out_in_use + sum(self.ws_bytes)

    def pending(self):
        return (len(self.queue) + len(self.ready) + len(self.ready_large) +
                sum(len(s[1]) for s in self.inflight if s is not None))

    def drain(self):
        while self.pending():
            self.pump()


def poisson_arrivals(rng, n, qps):
    return np.cumsum(rng.exponential(1.0 / qps, n))


def run_served(server, requests, clock=time.perf_counter):
    """Feed `requests` (with arrival offsets in seconds) in real time."""
    t0 = clock()
    i = 0
    order = sorted(requests, key=lambda r: r.arrival)
    while i < len(order) or server.pending():
        now = clock() - t0
        while i < len(order) and order[i].arrival <= now:
            server.submit(order[i])
            i += 1
        server.pump()
    return t0


def run_plain(backend, requests, stream, clock=time.perf_counter):
    """Baseline: each request launched alone at its arrival on one stream."""
    t0 = clock()
    order = sorted(requests, key=lambda r: r.arrival)
    inflight = collections.deque()
    i = 0
    while i < len(order) or inflight:
        now = clock() - t0
        while i < len(order) and order[i].arrival <= now:
            r = order[i]
            r.dst = backend.alloc(r.out_shape())
            r.admitted = r.launched = clock()
            r.run_alone(r.dst, stream)
            inflight.append((backend.record(stream), r))
            i += 1
        while inflight and backend.query(inflight[0][0]):
            inflight.popleft()[1].done = clock()
    return t0


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
