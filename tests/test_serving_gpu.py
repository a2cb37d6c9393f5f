"""Served results are the same alone and in the mix.

A stream of mixed requests (every served op, uneven sizes and channel
counts) goes through the continuous-batching server; every request's
output must equal the same op run alone through its single-image entry
point, byte for byte, and a sample must equal the CPU oracle.
"""

import numpy as np
import pytest

import kleidicv_hip as K
import oracle_lib as O
import serving as S
import workloads as WL

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def assert_same(got, want, what=""):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    if not np.array_equal(got, want):
        bad = np.argwhere(got != want)
        i = tuple(bad[0])
        raise AssertionError("%s: %d mismatches, first at %s: got %d want %d" %
                             (what, len(bad), i, got[i], want[i]))


def oracle(r):
    src = r.src.cpu().numpy()
    if r.op == "gaussian5":
        return O.gaussian_blur(src, 5, 0.0, "REVERSE")[1]
    if r.op == "median5":
        return O.median_blur(src, 5, "REPLICATE")[1]
    if r.op == "resize_half":
        h, w, _ = r.out_shape()
        return O.resize_linear(src, w, h)[1]
    if r.op == "warp":
        h, w, _ = r.out_shape()
        return O.warp_perspective(src, r.T, w, h, "LINEAR", "REPLICATE")[1]
    return O.yuv_sp_to_rgb(src, r.uv.cpu().numpy(), K.NV12_TO_RGB)[1]


@pytest.mark.parametrize("slots,large_px,large_batch,qps",
                         [(1, 1 << 18, 1, 1e9), (2, 1 << 18, 1, 1e9), (3, 1 << 15, 1, 1e9),
                          (2, 1 << 15, 8, 1e9), (2, 1 << 15, 8, 30000.0)])
def test_served_results_equal_alone_results(slots, large_px, large_batch, qps):
    dev = torch.device("cuda:0")
    corpus = WL.frame_corpus(torch, dev, n=10, seed=31, lo=(64, 64), hi=(480, 360))
    torch.cuda.synchronize()
    reqs = WL.serving_requests(corpus, 150, qps=qps, seed=32 + slots + large_batch)
    srv = S.Server(torch, dev, hbm_budget=1 << 30, slots=slots,
                   max_batch_ops=24, workspace_reserve=16 << 20, large_px=large_px,
                   large_batch=large_batch)
    S.run_served(srv, reqs)
    torch.cuda.synchronize()
    assert len(srv.completed) == len(reqs)
    if qps > 1e6:  # a burst: requests share batch launches
        assert srv.batches > 1 and max(sum(1 for r in reqs if r.batch == b)
                                       for b in range(1, srv.batches + 1)) > 1
    stream = torch.cuda.Stream()
    alone = []
    for r in reqs:
        alone.append(r.run_alone(torch.empty(r.out_shape(), dtype=torch.uint8, device=dev),
                                 stream))
    torch.cuda.synchronize()
    for r, a in zip(reqs, alone):
        assert_same(r.dst.cpu().numpy(), a.cpu().numpy(), "request %d (%s)" % (r.rid, r.op))
    seen = set()
    for r in reqs:
        if r.op in seen:
            continue
        seen.add(r.op)
        want = oracle(r)
        got = r.dst.cpu().numpy()
        assert_same(got.reshape(want.shape), want, "oracle %s" % r.op)


def test_plain_baseline_runs_the_same_requests():
    dev = torch.device("cuda:0")
    corpus = WL.frame_corpus(torch, dev, n=6, seed=41, lo=(64, 64), hi=(320, 240))
    torch.cuda.synchronize()
    reqs = WL.serving_requests(corpus, 40, qps=50000.0, seed=42)
    t0 = S.run_plain(torch, dev, reqs, hbm_budget=1 << 30)
    torch.cuda.synchronize()
    stats = S.latency_stats(reqs, t0)
    assert stats["requests"] == 40 and stats["p99_ms"] >= stats["p50_ms"] > 0
    assert all(r.alone and r.slot == 0 for r in reqs)
    assert sorted(r.batch for r in reqs) == list(range(1, 41))
    for r in reqs[:10]:
        want = oracle(r)
        assert_same(r.dst.cpu().numpy().reshape(want.shape), want, "plain %s" % r.op)


def test_invalid_request_fails_alone_and_the_rest_are_served():
    # a median of even size is declined by validation (NOT_IMPLEMENTED, as
    # the single-image entry point declines it); it completes as failed
    # whether it is launched alone or inside a batch, and the others run
    dev = torch.device("cuda:0")
    corpus = WL.frame_corpus(torch, dev, n=4, seed=61, lo=(64, 64), hi=(200, 160))
    torch.cuda.synchronize()
    reqs = WL.serving_requests(corpus, 30, qps=1e9, seed=62)

    class Bad(S.Request):
        __slots__ = ()

        def desc(self):
            d = super().desc()
            d.kernel_size = 4
            return d

    f = corpus[0]
    bad = [Bad(100 + i, "median5", f["src"], T=f["T"], arrival=0.0) for i in range(2)]
    srv = S.Server(torch, dev, hbm_budget=1 << 30, slots=2, max_batch_ops=8,
                   workspace_reserve=16 << 20)
    S.run_served(srv, bad[:1] + reqs + bad[1:])
    torch.cuda.synchronize()
    assert all(b.error == K.NOT_IMPLEMENTED for b in bad)
    assert all(r.error is None for r in reqs)
    stream = torch.cuda.Stream()
    for r in reqs:
        a = r.run_alone(torch.empty(r.out_shape(), dtype=torch.uint8, device=dev), stream)
        torch.cuda.synchronize()
        assert_same(r.dst.cpu().numpy(), a.cpu().numpy(), "request %d (%s)" % (r.rid, r.op))
