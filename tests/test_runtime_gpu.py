"""Runtime behaviour of the C ABI on the device: ordering of the drop-in
entry points after pending device work, and allocation failures.

The drop-in `kleidicv_*` functions behave like the reference's synchronous
CPU calls, so a source written by device work that is still queued must be
read only after that work.  Allocation failures are forced through the
library's test hook, as the reference forces malloc failures
(test/framework/wrap_malloc.cpp, e.g. test_resize_linear.cpp:844-867): the
call returns KLEIDICV_ERROR_ALLOCATION, and a batch whose rebuild fails keeps
running its previous plan.
"""

import numpy as np
import pytest

import kleidicv_hip as K
import oracle_lib as O
import workloads as WL

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev():
    return torch.device("cuda:0")


def assert_same(got, want, what=""):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    if not np.array_equal(got, want):
        bad = np.argwhere(got != want)
        i = tuple(bad[0])
        raise AssertionError("%s: %d mismatches, first at %s: got %d want %d" %
                             (what, len(bad), i, got[i], want[i]))


@pytest.fixture(autouse=True)
def disarm():
    yield
    K.fail_device_allocation(-1)


def test_drop_in_reads_source_after_pending_device_work():
    rng = np.random.default_rng(71)
    want_src = rng.integers(0, 256, (720, 1280), dtype=np.uint8)
    _, want = O.gaussian_blur(want_src, 5, 0.0, "REVERSE")
    staged = torch.from_numpy(want_src).to(dev())
    for rep in range(3):
        src = torch.zeros((720, 1280), dtype=torch.uint8, device=dev())
        dst = torch.zeros_like(src)
        torch.cuda.synchronize()
        # the source is written by a copy queued behind a long busy-wait on
        # torch's stream; the drop-in call is made at once, from the host
        torch.cuda._sleep(50_000_000)
        src.copy_(staged)
        K.gaussian_blur(src, 5, 0.0, "REVERSE", dst=dst)  # drop-in, device pointers
        # the result is in dst when the call returns: read it on another stream
        side = torch.cuda.Stream()
        with torch.cuda.stream(side):
            got = dst.cpu().numpy()
        assert_same(got, want, "rep %d" % rep)
    torch.cuda.synchronize()


def test_allocation_failure_returns_allocation_error():
    rng = np.random.default_rng(72)
    src = rng.integers(0, 256, (90, 130, 3), dtype=np.uint8)
    _, want = O.resize_linear(src, 171, 121)  # generic path: coordinate tables
    d_src = torch.from_numpy(src).to(dev())
    d_dst = torch.empty((121, 171, 3), dtype=torch.uint8, device=dev())
    torch.cuda.synchronize()
    # device API on a stream the library has not seen (its table buffer is
    # allocated on first use): the tables allocation fails
    fresh = torch.cuda.Stream()
    K.fail_device_allocation(0)
    with pytest.raises(K.KleidiCVError) as e:
        K.resize_linear(d_src, 171, 121, dst=d_dst, stream=fresh, device_api=True)
    assert e.value.code == K.ALLOCATION
    # drop-in on host buffers: each of the staging and table allocations in
    # turn (the library's staging buffer dropped first, so the call allocates)
    for after in range(3):
        K.release_staging()
        K.fail_device_allocation(after)
        with pytest.raises(K.KleidiCVError) as e:
            K.resize_linear(src, 171, 121)
        assert e.value.code == K.ALLOCATION, after
    K.fail_device_allocation(-1)
    assert_same(K.resize_linear(src, 171, 121), want, "after the failures")
    K.resize_linear(d_src, 171, 121, dst=d_dst, stream=fresh, device_api=True)
    torch.cuda.synchronize()
    assert_same(d_dst.cpu().numpy(), want, "device API after the failures")


def test_device_resize_reuses_its_stream_table_buffer():
    # calls on one stream share a table buffer that grows on demand; back to
    # back calls of different sizes (tables written and read in stream order)
    # and calls on a second stream give the oracle's bytes
    rng = np.random.default_rng(74)
    cases = [((90, 130, 3), 171, 121), ((300, 500, 1), 777, 411), ((60, 50, 4), 37, 29),
             ((300, 500, 1), 200, 120), ((90, 130, 3), 171, 121)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    jobs = []
    for i, (shape, w, h) in enumerate(cases * 2):
        src = rng.integers(0, 256, shape, dtype=np.uint8)
        d_src = torch.from_numpy(src).to(dev())
        d_dst = torch.empty((h, w, shape[2]), dtype=torch.uint8, device=dev())
        jobs.append((src, d_src, d_dst, w, h, streams[i % 2]))
    torch.cuda.synchronize()
    for _, d_src, d_dst, w, h, st in jobs:
        K.resize_linear(d_src, w, h, dst=d_dst, stream=st, device_api=True)
    torch.cuda.synchronize()
    for src, _, d_dst, w, h, _ in jobs:
        _, want = O.resize_linear(src, w, h)
        assert_same(d_dst.cpu().numpy(), want, "%s -> %dx%d" % (src.shape, w, h))


def test_failed_batch_rebuild_keeps_the_previous_plan():
    recs = WL.pyramid_batch(torch, dev(), n=8, seed=73, lo=64, hi=300)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    first = K.Tuning(4, 0, 1, 0)
    with K.Batch(WL.pyramid_descs(recs), tuning=first, keep=recs) as b:
        items = b.work_items
        stream.wait_stream(torch.cuda.current_stream())
        b.run(stream)
        torch.cuda.synchronize()
        want = [t.clone() for r in recs for t in [r["gauss"]] + r["levels"]]
        K.fail_device_allocation(0)
        with pytest.raises(K.KleidiCVError) as e:
            b.tuning = K.Tuning(1, 16, 0, 0)
        assert e.value.code == K.ALLOCATION
        assert b.tuning.as_dict() == first.as_dict()
        assert b.work_items == items
        for r in recs:
            r["gauss"].fill_(0xA5)
            for lv in r["levels"]:
                lv.fill_(0x5A)
        stream.wait_stream(torch.cuda.current_stream())
        b.run(stream)
        torch.cuda.synchronize()
        got = [t for r in recs for t in [r["gauss"]] + r["levels"]]
        for i, (g, w) in enumerate(zip(got, want)):
            assert torch.equal(g, w), "output %d after the failed rebuild" % i
