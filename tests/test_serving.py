"""Scheduling logic of the serving engine, on the host.

The native engine (kleidicv_hip_server_*) runs against its host stand-in,
whose "device" work completes a set number of polls after launch, so
admission, batching and completion order are checked without a GPU.
"""

import numpy as np
import pytest

import kleidicv_hip as K
import serving as S


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1.0
        return self.t


def standin(polls, **kw):
    """The native engine on its host stand-in (no GPU); outputs return to
    the ring at completion unless a test says otherwise."""
    kw.setdefault("clock", Clock())
    kw.setdefault("release_done", True)
    return S.Server(standin_polls=polls, **kw)


def launches(srv):
    """Launches in launch order, rebuilt from the completed requests: a
    batch as its request ids, a lone launch as ["alone", id]."""
    by = {}
    for r in srv.completed:
        by.setdefault(r.batch, []).append(r)
    out = []
    for b in sorted(by):
        rs = sorted(by[b], key=lambda r: r.sid)
        out.append(["alone", rs[0].rid] if rs[0].alone else [r.rid for r in rs])
    return out


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
    srv = standin(2, hbm_budget=1 << 30, slots=2, max_batch_ops=4)
    reqs = [req(i, S.OPS[i % len(S.OPS)]) for i in range(23)]
    for r in reqs:
        srv.submit(r)
    srv.drain()
    assert sorted(r.rid for r in srv.completed) == list(range(23))
    adm = [r.admitted for r in reqs]
    assert adm == sorted(adm)                     # FIFO: never admitted before an earlier one
    ls = launches(srv)
    assert all(len(b) <= 4 for b in ls if b[0] != "alone")
    assert sum(1 if b[0] == "alone" else len(b) for b in ls) == 23
    assert all(r.admitted <= r.launched < r.done for r in reqs)
    assert srv.out_in_use == 0 and srv.in_use == srv.ws_bytes == 2 * (64 << 20)
    assert srv.pending() == 0


def test_budget_bounds_resident_bytes_and_blocks_the_queue_head():
    one = 64 * 64
    reserve = 4096   # work lists of up to 4 ops per slot
    # outputs of 3 requests fit beside the two reserves; the 4th waits
    budget = 3 * one + 2 * reserve
    srv = standin(3, hbm_budget=budget, slots=2, max_batch_ops=8, workspace_reserve=reserve)
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
    assert all(len(b) <= 3 for b in launches(srv) if b[0] != "alone")
    big = standin(3, hbm_budget=3 * 4096 - 1, workspace_reserve=4096)
    with pytest.raises(K.KleidiCVError) as e:
        big.submit(req(0))   # a 4096-byte output does not fit beside two reserves
    assert e.value.code == K.RANGE
    with pytest.raises(K.KleidiCVError):   # the reserves leave nothing for outputs
        standin(3, hbm_budget=2 * 4096, workspace_reserve=4096)


def test_batches_larger_than_the_workspace_reserve_are_split():
    # the stand-in's work lists take 1024 bytes per op
    srv = standin(2, hbm_budget=1 << 30, slots=1, max_batch_ops=64, workspace_reserve=3 * 1024)
    for i in range(10):
        srv.submit(req(i))
    srv.drain()
    assert sorted(r.rid for r in srv.completed) == list(range(10))
    ls = launches(srv)
    assert all(len(b) <= 3 for b in ls)
    assert [rid for b in ls for rid in (b[1:] if b[0] == "alone" else b)] == list(range(10))


def test_arrivals_during_a_batch_launch_while_it_runs():
    srv = standin(6, hbm_budget=1 << 30, slots=2, max_batch_ops=64)
    first = [req(i) for i in range(5)]
    for r in first:
        srv.submit(r)
    srv.pump()                       # batch 1 launched
    assert srv.batches == 1
    late = [req(100 + i) for i in range(3)]
    for r in late:
        srv.submit(r)
    srv.pump()                       # batch 2 launched while batch 1 runs
    assert srv.batches == 2
    assert all(r.done is None for r in first)
    srv.drain()
    assert launches(srv) == [[0, 1, 2, 3, 4], [100, 101, 102]]
    assert max(r.launched for r in late) < min(r.done for r in first)


def test_single_slot_of_depth_one_serialises_batches():
    srv = standin(4, hbm_budget=1 << 30, slots=1, depth=1)
    for i in range(3):
        srv.submit(req(i))
    srv.pump()
    srv.submit(req(9))
    srv.pump()
    assert srv.batches == 1           # the only slot is busy
    srv.drain()
    assert launches(srv) == [[0, 1, 2], ["alone", 9]]   # a lone request runs alone


def test_large_requests_run_alone_on_slot_zero_and_small_ones_pass_them():
    srv = standin(6, hbm_budget=1 << 30, slots=2, max_batch_ops=8, large_px=64 * 64, depth=1)
    big = [req(i, h=64, w=64) for i in range(3)]           # 4096 px: large
    small = [req(10 + i, h=16, w=16) for i in range(5)]    # small
    for r in big[:2] + small + big[2:]:
        srv.submit(r)
    srv.pump()
    srv.drain()
    ls = launches(srv)
    # slot 0 took the first large request alone, slot 1 the small batch
    assert ls[0] == ["alone", 0] and big[0].slot == 0
    assert ls[1] == [10, 11, 12, 13, 14] and small[0].slot == 1
    alone = [b[1] for b in ls if b[0] == "alone"]
    assert alone == [0, 1, 2]                              # large: arrival order, one at a time
    assert all(r.slot == 0 for r in big)
    batched = [rid for b in ls if b[0] != "alone" for rid in b]
    assert sorted(batched) == [10, 11, 12, 13, 14]
    assert max(r.done for r in small) < min(r.done for r in big[1:])
    assert sorted(r.rid for r in srv.completed) == [0, 1, 2, 10, 11, 12, 13, 14]


def test_backlogged_large_requests_share_batch_launches():
    srv = standin(4, hbm_budget=1 << 30, slots=2, max_batch_ops=8, large_px=64 * 64, depth=1,
                  large_batch=3)
    big = [req(i, h=64, w=64) for i in range(6)]
    for r in big[:5]:
        srv.submit(r)
    srv.pump()
    assert srv.batches == 1                    # backlog: one batch of large_batch
    srv.drain()
    assert launches(srv) == [[0, 1, 2], [3, 4]]  # the rest of the backlog, still in order
    srv.submit(big[5])
    srv.drain()
    assert launches(srv)[-1] == ["alone", 5]     # no backlog: alone again
    assert [r.rid for r in srv.completed] == list(range(6))
    with pytest.raises(K.KleidiCVError):
        standin(4, large_batch=0)


def test_a_slot_queues_up_to_depth_launches():
    srv = standin(50, hbm_budget=1 << 30, slots=1, max_batch_ops=2, depth=3)
    reqs = [req(i) for i in range(10)]
    for r in reqs:
        srv.submit(r)
    srv.pump()
    assert srv.batches == 3                            # three queued, none retired yet
    assert srv.pending() == 10
    srv.drain()
    assert launches(srv) == [[0, 1], [2, 3], [4, 5], [6, 7], [8, 9]]
    assert [r.rid for r in srv.completed] == list(range(10))   # retired in launch order


def test_output_ring_returns_space_oldest_first():
    # 16x16 outputs are one 256-byte ring block each; the ring holds four
    srv = standin(1, hbm_budget=1 << 30, slots=1, max_batch_ops=1, ring_bytes=1024,
                  release_done=False)
    reqs = [req(i, h=16, w=16) for i in range(6)]
    for r in reqs:
        srv.submit(r)
    with pytest.raises(RuntimeError):   # 4 and 5 wait for ring space held by 0..3
        srv.drain()
    assert [r.rid for r in srv.completed] == [0, 1, 2, 3]
    assert srv.pending() == 2 and srv.counters().stalled
    srv.release(reqs[1])            # out of order: nothing returns yet
    for _ in range(4):
        srv.pump()
    assert len(srv.completed) == 4
    srv.release(reqs[0])            # blocks 0 and 1 return: room for one block
    with pytest.raises(RuntimeError):   # (a wrapped head never meets the tail)
        srv.drain()
    assert [r.rid for r in srv.completed] == [0, 1, 2, 3, 4]
    srv.release(reqs[2])
    srv.drain()
    assert [r.rid for r in srv.completed] == list(range(6))
    with pytest.raises(K.KleidiCVError):
        srv.release(reqs[0])        # already returned and forgotten
    with pytest.raises(K.KleidiCVError) as e:
        standin(1, ring_bytes=1024).submit(req(7, h=64, w=64))  # larger than the whole ring
    assert e.value.code == K.RANGE


def test_admission_waits_for_output_buffers():
    # a ring of two 64x64 outputs, returned at completion
    srv = standin(3, hbm_budget=1 << 30, slots=2, max_batch_ops=4, ring_bytes=2 * 4096,
                  release_done=True)
    for i in range(7):
        srv.submit(req(i))
    peak = 0
    while srv.pending():
        srv.pump()
        peak = max(peak, srv.out_in_use)
    assert peak == 2 * 4096
    assert sorted(r.rid for r in srv.completed) == list(range(7))


def test_poisson_arrivals_rate():
    rng = np.random.default_rng(1)
    t = S.poisson_arrivals(rng, 20000, 5000.0)
    assert np.all(np.diff(t) > 0)
    assert abs(20000 / t[-1] - 5000.0) / 5000.0 < 0.03
