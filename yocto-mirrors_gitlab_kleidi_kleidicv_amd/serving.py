"""Serving loop: continuous batching of image requests on one GPU.

Requests (an op on a frame that is already in device memory) arrive at any
time.  The server keeps every buffer resident on the device and admits a
request once its output fits the HBM budget.  Admitted requests are split by
size.  A large request (at least `large_px` output pixels) fills the GPU on
its own, so it is launched alone through its single-image device entry point
on the large-request slot, in arrival order; when more large requests back
up than the slot queues (the device is busy and the host sets the pace), up
to `large_batch` of them share one batch launch.  Small requests are packed into
batches (kleidicv_hip_batch_launch: one persistent launch per kernel variant)
on the other stream slots, so they never wait behind a large one and the
launch cost is shared (a lone small request also runs through its own entry
point); the large-request slot also takes small batches when no large
request waits.  Each slot keeps up to `depth` launches queued on
its stream, so the device never waits for the host between them, and
batches on different slots overlap on the device.  A request completes
when the host observes its launch's event; its output is then handed to the
caller and its budget released.

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
                 "done", "batch", "_shape")

    def __init__(self, rid, op, src, uv=None, T=None, arrival=0.0):
        if op not in OPS:
            raise ValueError("unknown op %r" % op)
        self.rid, self.op, self.src, self.uv, self.T = rid, op, src, uv, T
        self.arrival = arrival
        self.dst = None
        self.admitted = self.launched = self.done = None
        self.batch = None
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


class DeviceArena:
    """Byte ranges of one device allocation handed out as a ring.

    Blocks are taken at the head in allocation order and may be released in
    any order; the space of the oldest blocks returns once they and every
    block before them are released.  Requests complete in roughly arrival
    order, so the ring stays compact, and taking a block costs no device
    allocation.
    """

    def __init__(self, size, align=256):
        self.size, self.align = int(size), align
        self.head = 0
        self.blocks = collections.deque()   # [offset, nbytes, released]

    def take(self, nbytes):
        """Offset of a new block, or None while the ring is full."""
        n = -(-int(nbytes) // self.align) * self.align
        if n > self.size:
            raise ValueError("block of %d bytes exceeds the arena (%d bytes)" % (n, self.size))
        if not self.blocks:
            self.head = 0
        tail = self.blocks[0][0] if self.blocks else 0
        if not self.blocks or self.head > tail:     # free: [head, size) and [0, tail)
            if self.size - self.head >= n:
                off = self.head
            elif tail > n:                           # wrap; the end of the ring idles
                off = 0
            else:
                return None
        elif tail - self.head > n:                   # wrapped; free: [head, tail)
            off = self.head
        else:
            return None
        block = [off, n, False]
        self.blocks.append(block)
        self.head = off + n
        return block

    def release(self, block):
        block[2] = True
        while self.blocks and self.blocks[0][2]:
            self.blocks.popleft()


class TorchBackend:
    """Device side through torch tensors, streams and events.

    With arena_bytes, output buffers are views into one device allocation
    managed as a DeviceArena (no allocator call per request); alloc returns
    None while it is full and free returns a buffer to it.  Without, each
    output is a torch allocation and free does nothing.
    """

    def __init__(self, torch, device, arena_bytes=None):
        self.torch, self.device = torch, device
        self.arena = DeviceArena(arena_bytes) if arena_bytes else None
        self.arena_buf = None
        self.blocks = {}

    def stream(self):
        return self.torch.cuda.Stream(device=self.device)

    def alloc(self, shape):
        if self.arena is None:
            return self.torch.empty(shape, dtype=self.torch.uint8, device=self.device)
        if self.arena_buf is None:
            self.arena_buf = self.torch.empty(self.arena.size, dtype=self.torch.uint8,
                                              device=self.device)
        n = 1
        for d in shape:
            n *= d
        block = self.arena.take(n)
        if block is None:
            return None
        t = self.arena_buf[block[0]:block[0] + n].view(shape)
        self.blocks[t.data_ptr()] = block
        return t

    def free(self, tensor):
        if self.arena is not None:
            self.arena.release(self.blocks.pop(tensor.data_ptr()))

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
    large-request slot; with one slot, everything runs on it.  depth bounds
    the launches queued on one slot's stream.  A slot's work lists are
    uploaded on its own stream, so a queued batch can reuse the workspace;
    a replaced workspace stays referenced until the launches that read it
    retire.
    """

    def __init__(self, backend, hbm_budget=64 << 30, slots=2, max_batch_ops=512,
                 workspace_reserve=64 << 20, tuning=None, clock=time.perf_counter,
                 large_px=1 << 18, depth=4, release_done=False, large_batch=1):
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
        if depth < 1:
            raise ValueError("depth must be >= 1")
        self.depth = depth
        self.release_done = release_done
        self.streams = [backend.stream() for _ in range(slots)]
        # per slot, in launch order: (event, requests, workspace read)
        self.inflight = [collections.deque() for _ in range(slots)]
        self.ws = [None] * slots
        self.ws_bytes = [0] * slots
        self.large_px = large_px
        if large_batch < 1:
            raise ValueError("large_batch must be >= 1")
        self.large_batch = large_batch
        self.queue = collections.deque()   # arrived, waiting for budget
        self.ready = collections.deque()   # admitted small requests, waiting for a slot
        self.ready_large = collections.deque()  # admitted large requests
        self.completed = []
        self.batches = 0

    # -- admission ----------------------------------------------------------
    def submit(self, req):
        self.queue.append(req)

    def _admit(self, now):
        # Admitted requests hold their output buffers; admit no more than the
        # slots can queue, so allocation interleaves with launches.
        cap = self.depth * self.max_batch_ops * len(self.streams)
        while self.queue and len(self.ready) + len(self.ready_large) < cap:
            req = self.queue[0]
            need = req.out_bytes()
            if need > self.out_budget:
                raise ValueError("request %d needs %d bytes, more than the whole budget"
                                 % (req.rid, need))
            if self.out_in_use + need > self.out_budget:
                return  # FIFO: later requests wait behind the head
            dst = self.be.alloc(req.out_shape())
            if dst is None:
                return  # output buffers exhausted until earlier requests complete
            self.queue.popleft()
            self.out_in_use += need
            self.peak = max(self.peak, self.in_use)
            req.dst = dst
            req.admitted = now
            if req.out_pixels() >= self.large_px:
                self.ready_large.append(req)
            else:
                self.ready.append(req)

    # -- completion -----------------------------------------------------------
    def _retire(self, now):
        for q in self.inflight:
            while q and self.be.query(q[0][0]):
                _, reqs, _ = q.popleft()
                for r in reqs:
                    r.done = now
                    self.out_in_use -= r.out_bytes()
                    self.completed.append(r)
                    if self.release_done:
                        self.be.free(r.dst)

    # -- batching -----------------------------------------------------------
    def _launch(self, now):
        for i in range(len(self.streams)):
            while len(self.inflight[i]) < self.depth and self._launch_one(i, now):
                pass

    def _launch_one(self, i, now):
        """Launches one request or batch on slot i; False when nothing fits."""
        if i == 0 and self.ready_large:
            # Large requests go in arrival order, one per launch; with a
            # backlog of more than `depth` (more than the slot's queue
            # holds: the host sets the pace) up to large_batch of them share
            # one batch launch.  A batched request completes with its whole
            # batch, so below that backlog they stay alone for latency.
            if self.large_batch == 1 or len(self.ready_large) <= self.depth:
                req = self.ready_large.popleft()
                self.be.launch_one(req, self.streams[0])
                self._started(0, [req], None, now)
                return True
            self._launch_batch(0, self.ready_large, self.large_batch, now)
            return True
        if not self.ready:
            return False
        if len(self.ready) == 1:
            # a lone small request: its own entry point is cheaper than a batch
            req = self.ready.popleft()
            self.be.launch_one(req, self.streams[i])
            self._started(i, [req], None, now)
            return True
        self._launch_batch(i, self.ready, self.max_batch_ops, now)
        return True

    def _launch_batch(self, i, queue, limit, now):
        """One batch launch on slot i of up to `limit` requests from `queue`."""
        reqs = []
        while queue and len(reqs) < limit:
            reqs.append(queue.popleft())
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
            queue.extendleft(reversed(reqs[keep:]))
            reqs, descs = reqs[:keep], descs[:keep]
        self._started(i, reqs, self.ws[i], now)

    def _started(self, i, reqs, workspace, now):
        ev = self.be.record(self.streams[i])
        self.batches += 1
        for r in reqs:
            r.launched = now
            r.batch = self.batches
        self.inflight[i].append((ev, reqs, workspace))

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
                sum(len(e[1]) for q in self.inflight for e in q))

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


def run_plain(backend, requests, stream, clock=time.perf_counter, release_done=False):
    """Baseline: each request launched alone at its arrival on one stream.
    A request whose output buffer is not available waits (arena backend)."""
    t0 = clock()
    order = sorted(requests, key=lambda r: r.arrival)
    inflight = collections.deque()
    i = 0
    while i < len(order) or inflight:
        now = clock() - t0
        while i < len(order) and order[i].arrival <= now:
            r = order[i]
            r.dst = backend.alloc(r.out_shape())
            if r.dst is None:
                break
            r.admitted = r.launched = clock()
            r.run_alone(r.dst, stream)
            inflight.append((backend.record(stream), r))
            i += 1
        while inflight and backend.query(inflight[0][0]):
            r = inflight.popleft()[1]
            r.done = clock()
            if release_done:
                backend.free(r.dst)
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
