"""Scheduling logic of the serving loop, on the host.

The server runs against a host-side backend whose "device" work completes a
set number of polls after launch, so admission, batching and completion
order can be checked without a GPU.
"""

import numpy as np
import pytest

import serving as S


class FakeEvent:
    def __init__(self, polls):
        self.left = polls


class FakeBackend:
    def __init__(self, polls=3):
        self.polls = polls
        self.launches = []

    def stream(self):
        return object()

    def alloc(self, shape):
        return np.zeros(shape, np.uint8)

    def workspace(self, nbytes):
        return np.zeros(max(int(nbytes), 1), np.uint8)

    def launch(self, descs, workspace, stream, tuning, requests):
        assert len(descs) == len(requests)
        need = 1024 * len(descs)
        if workspace is None or workspace.size < need:
            return need
        self.launches.append([r.rid for r in requests])
        return 0

    def launch_one(self, request, stream):
        self.launches.append(["alone", request.rid])

    def record(self, stream):
        return FakeEvent(self.polls)

    def query(self, ev):
        ev.left -= 1
        return ev.left <= 0


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1.0
        return self.t


def frame(h, w, c=1):
    return np.zeros((h, w, c), np.uint8)


def req(rid, op="gaussian5", h=64, w=64, c=1):
    if op == "nv12":
        return S.Request(rid, op, np.zeros((h, w), np.uint8), uv=np.zeros((h // 2, w), np.uint8))
    return S.Request(rid, op, frame(h, w, c), T=np.eye(3, dtype=np.float32))


def test_output_shapes():
    assert req(0, "resize_half", 65, 131, 3).out_shape() == (33, 66, 3)
    assert req(1, "nv12", 64, 96).out_shape() == (64, 96, 3)
    assert req(2, "median5", 70, 80, 4).out_shape() == (70, 80, 4)
    with pytest.raises(ValueError):
        S.Request(3, "sobel", frame(8, 8))


def test_every_request_completes_once_in_fifo_admission_order():
    be = FakeBackend(polls=2)
    srv = S.Server(be, hbm_budget=1 << 30, slots=2, max_batch_ops=4, clock=Clock())
    reqs = [req(i, S.OPS[i % len(S.OPS)]) for i in range(23)]
    for r in reqs:
        srv.submit(r)
    srv.drain()
    assert sorted(r.rid for r in srv.completed) == list(range(23))
    admitted = sorted(reqs, key=lambda r: r.admitted)
    assert [r.rid for r in admitted] == list(range(23))
    assert all(len(b) <= 4 for b in be.launches)
    assert sum(len(b) for b in be.launches) == 23
    assert all(r.admitted <= r.launched < r.done for r in reqs)
    assert srv.out_in_use == 0 and srv.in_use == sum(srv.ws_bytes)


def test_budget_bounds_resident_bytes_and_blocks_the_queue_head():
    be = FakeBackend(polls=3)
    one = 64 * 64
    reserve = 4096   # work lists of up to 4 ops per slot
    # outputs of 3 requests fit beside the two reserves; the 4th waits
    budget = 3 * one + 2 * reserve
    srv = S.Server(be, hbm_budget=budget, slots=2, max_batch_ops=8,
                   workspace_reserve=reserve, clock=Clock())
    for i in range(10):
        srv.submit(req(i))
    peak_seen = max_out = 0
    while srv.pending():
        srv.pump()
        peak_seen = max(peak_seen, srv.in_use)
        max_out = max(max_out, srv.out_in_use)
        assert srv.in_use <= budget
        assert srv.out_in_use <= 3 * one
    assert srv.peak <= budget and peak_seen <= budget and max_out == 3 * one
    assert len(srv.completed) == 10
    assert all(len(b) <= 3 for b in be.launches)
    with pytest.raises(ValueError):
        big = S.Server(FakeBackend(), hbm_budget=3 * 4096 - 1, workspace_reserve=4096,
                       clock=Clock())
        big.submit(req(0))   # a 4096-byte output does not fit beside two reserves
        big.pump()


def test_batches_larger_than_the_workspace_reserve_are_split():
    be = FakeBackend(polls=2)
    srv = S.Server(be, hbm_budget=1 << 30, slots=1, max_batch_ops=64,
                   workspace_reserve=3 * 1024, clock=Clock())
    for i in range(10):
        srv.submit(req(i))
    srv.drain()
    assert sorted(r.rid for r in srv.completed) == list(range(10))
    assert all(len(b) <= 3 for b in be.launches)
    assert [rid for b in be.launches for rid in b] == list(range(10))


def test_arrivals_during_a_batch_overlap_it_on_another_slot():
    be = FakeBackend(polls=6)
    clock = Clock()
    srv = S.Server(be, hbm_budget=1 << 30, slots=2, max_batch_ops=64, clock=clock)
    first = [req(i) for i in range(5)]
    for r in first:
        srv.submit(r)
    srv.pump()                       # batch 1 launched
    assert len(be.launches) == 1
    late = [req(100 + i) for i in range(3)]
    for r in late:
        srv.submit(r)
    srv.pump()                       # batch 2 launched while batch 1 runs
    assert len(be.launches) == 2 and be.launches[1] == [100, 101, 102]
    assert all(r.done is None for r in first)
    srv.drain()
    assert max(r.launched for r in late) < min(r.done for r in first)


def test_single_slot_of_depth_one_serialises_batches():
    be = FakeBackend(polls=4)
    srv = S.Server(be, hbm_budget=1 << 30, slots=1, clock=Clock(), depth=1)
    for i in range(3):
        srv.submit(req(i))
    srv.pump()
    srv.submit(req(9))
    srv.pump()
    assert len(be.launches) == 1      # the only slot is busy
    srv.drain()
    assert be.launches == [[0, 1, 2], ["alone", 9]]   # a lone request runs alone


def test_large_requests_run_alone_on_slot_zero_and_small_ones_pass_them():
    be = FakeBackend(polls=6)
    srv = S.Server(be, hbm_budget=1 << 30, slots=2, max_batch_ops=8, clock=Clock(),
                   large_px=64 * 64, depth=1)
    big = [req(i, h=64, w=64) for i in range(3)]           # 4096 px: large
    small = [req(10 + i, h=16, w=16) for i in range(5)]    # small
    for r in big[:2] + small + big[2:]:
        srv.submit(r)
    srv.pump()
    # slot 0 took the first large request alone, slot 1 the small batch
    assert be.launches[0] == ["alone", 0]
    assert be.launches[1] == [10, 11, 12, 13, 14]
    srv.drain()
    alone = [b[1] for b in be.launches if b[0] == "alone"]
    assert alone == [0, 1, 2]                              # large: arrival order, one at a time
    batched = [rid for b in be.launches if b[0] != "alone" for rid in b]
    assert sorted(batched) == [10, 11, 12, 13, 14]
    assert max(r.done for r in small) < min(r.done for r in big[1:])
    assert sorted(r.rid for r in srv.completed) == [0, 1, 2, 10, 11, 12, 13, 14]


def test_backlogged_large_requests_share_batch_launches():
    be = FakeBackend(polls=4)
    srv = S.Server(be, hbm_budget=1 << 30, slots=2, max_batch_ops=8, clock=Clock(),
                   large_px=64 * 64, depth=1, large_batch=3)
    big = [req(i, h=64, w=64) for i in range(6)]
    for r in big[:5]:
        srv.submit(r)
    srv.pump()
    assert be.launches == [[0, 1, 2]]          # backlog: one batch of large_batch
    srv.drain()
    assert be.launches == [[0, 1, 2], [3, 4]]  # the rest of the backlog, still in order
    srv.submit(big[5])
    srv.drain()
    assert be.launches[-1] == ["alone", 5]     # no backlog: alone again
    assert [r.rid for r in srv.completed] == list(range(6))
    with pytest.raises(ValueError):
        S.Server(be, large_batch=0)


def test_a_slot_queues_up_to_depth_launches():
    be = FakeBackend(polls=50)
    srv = S.Server(be, hbm_budget=1 << 30, slots=1, max_batch_ops=2, clock=Clock(), depth=3)
    for i in range(10):
        srv.submit(req(i))
    srv.pump()
    assert be.launches == [[0, 1], [2, 3], [4, 5]]     # three queued, none retired yet
    assert srv.pending() == 10
    srv.drain()
    assert [rid for b in be.launches for rid in b] == list(range(10))
    assert [r.rid for r in srv.completed] == list(range(10))   # retired in launch order


def test_device_arena_ring():
    a = S.DeviceArena(1024, align=64)
    b1, b2, b3 = a.take(300), a.take(300), a.take(300)   # 320 bytes each
    assert [b1[0], b2[0], b3[0]] == [0, 320, 640]
    assert a.take(200) is None                          # 64 bytes left at the end
    a.release(b2)                                       # out of order: nothing returns yet
    assert a.take(200) is None
    a.release(b1)                                       # [0, 640) returns
    b4 = a.take(500)                                    # wraps to the start
    assert b4[0] == 0
    assert a.take(200) is None                          # [512, 640) too small
    a.release(b3)
    a.release(b4)
    assert a.take(1024)[0] == 0                         # empty ring starts over
    with pytest.raises(ValueError):
        S.DeviceArena(128).take(129)


def test_admission_waits_for_output_buffers():
    class Limited(FakeBackend):
        def __init__(self):
            super().__init__(polls=3)
            self.live = 0
            self.peak = 0

        def alloc(self, shape):
            if self.live == 2:
                return None
            self.live += 1
            self.peak = max(self.peak, self.live)
            return np.zeros(shape, np.uint8)

        def free(self, t):
            self.live -= 1

    be = Limited()
    srv = S.Server(be, hbm_budget=1 << 30, slots=2, max_batch_ops=4, clock=Clock(),
                   release_done=True)
    for i in range(7):
        srv.submit(req(i))
    srv.drain()
    assert be.peak == 2 and be.live == 0
    assert sorted(r.rid for r in srv.completed) == list(range(7))


def test_poisson_arrivals_rate():
    rng = np.random.default_rng(1)
    t = S.poisson_arrivals(rng, 20000, 5000.0)
    assert np.all(np.diff(t) > 0)
    assert abs(20000 / t[-1] - 5000.0) / 5000.0 < 0.03
