"""Generates median5_net.h: the min/max network that selects the medians of
the eight 5x5 windows of a 2-column x 4-row output tile from its 6 x 8
footprint.

    python gen_median5_net.py        (writes median5_net.h beside this file)

Construction (exact selection, so every output equals the 5x5 median):
  - the windows of a group of outputs share a rectangular core; the core is
    sorted once, from sorted column (or row) strips merged by Batcher's
    odd-even merge;
  - the group is split in two along its longer side; each half adds the
    values its windows share beyond the parent's core (again sorted strips),
    merged into the parent's sorted core;
  - at every level only the ranks that can still be some window's median
    are kept: with n of a window's 25 values in the core and 25 - n outside,
    the core's value of rank r (0-based) is the window's value of rank r ..
    r + 25 - n, so ranks below 12 - (25 - n) are always below the median
    (dropped, the target rank shifting down) and ranks above 12 always
    above it (dropped);
  - a single window's median is then the target rank of the union of its
    kept core and its own remaining values.
The network is traced symbolically, operations whose results no output
needs are removed, and the result is checked against numpy's median on
random and low-entropy footprints before the header is written.
"""

import os
import sys

import numpy as np

K = 5
N = K * K
T = N // 2          # median rank, 0-based
TW, TH = 2, 4       # output tile
FW, FH = TW + K - 1, TH + K - 1


class Net:
    def __init__(self):
        self.ops = []            # (kind, a, b): kind 'min' | 'max'

    def mn(self, a, b):
        self.ops.append(("min", a, b))
        return ("op", len(self.ops) - 1)

    def mx(self, a, b):
        self.ops.append(("max", a, b))
        return ("op", len(self.ops) - 1)


def cmpex(net, a, b):
    return net.mn(a, b), net.mx(a, b)


SORT5 = [(0, 3), (1, 4), (0, 2), (1, 3), (0, 1), (2, 4), (1, 2), (3, 4), (2, 3)]
SMALL = {1: [], 2: [(0, 1)], 3: [(0, 2), (0, 1), (1, 2)],
         4: [(0, 1), (2, 3), (0, 2), (1, 3), (1, 2)], 5: SORT5}


def merge(net, A, B):
    """Batcher odd-even merge of sorted A and B (padding slots never compared)."""
    n = max(len(A), len(B))
    p = 1
    while p < n:
        p *= 2
    pos = list(A) + [None] * (p - len(A)) + list(B) + [None] * (p - len(B))

    def cmp(i, j):
        a, b = pos[i], pos[j]
        if a is not None and b is not None:
            pos[i], pos[j] = cmpex(net, a, b)
        elif a is None and b is not None:
            pos[i], pos[j] = b, None

    def rec(lo, n, r):
        m = 2 * r
        if m < n:
            rec(lo, n, m)
            rec(lo + r, n, m)
            i = lo + r
            while i + r < lo + n:
                cmp(i, i + r)
                i += m
        else:
            cmp(lo, lo + r)

    rec(0, 2 * p, 1)
    return [x for x in pos if x is not None]


def sort_small(net, ws):
    ws = list(ws)
    if len(ws) in SMALL:
        for a, b in SMALL[len(ws)]:
            ws[a], ws[b] = cmpex(net, ws[a], ws[b])
        return ws
    h = len(ws) // 2
    return merge(net, sort_small(net, ws[:h]), sort_small(net, ws[h:]))


def build():
    net = Net()
    cache = {}

    def strip(kind, idx, a, b):
        """Sorted values of footprint column idx rows a..b ('c') or row idx
        columns a..b ('r'), reusing the largest cached sub-strip."""
        key = (kind, idx, a, b)
        if key in cache:
            return cache[key]
        best = None
        for (k2, i2, a2, b2) in cache:
            if k2 == kind and i2 == idx and a <= a2 and b2 <= b and (a2, b2) != (a, b):
                if best is None or b2 - a2 > best[1] - best[0]:
                    best = (a2, b2)
        if best is None or best[1] == best[0]:
            cells = [("in", (y, idx)) if kind == "c" else ("in", (idx, y)) for y in range(a, b + 1)]
            if len(cells) <= 5:
                res = sort_small(net, cells)
            else:
                m = (a + b) // 2
                res = merge(net, strip(kind, idx, a, m), strip(kind, idx, m + 1, b))
        else:
            a2, b2 = best
            res = strip(kind, idx, a2, b2)
            if a < a2:
                res = merge(net, res, strip(kind, idx, a, a2 - 1))
            if b2 < b:
                res = merge(net, res, strip(kind, idx, b2 + 1, b))
        cache[key] = res
        return res

    def rect(group):
        return (max(o[0] for o in group), min(o[0] for o in group) + K - 1,
                max(o[1] for o in group), min(o[1] for o in group) + K - 1)

    def cells(r):
        y0, y1, x0, x1 = r
        return {(y, x) for y in range(y0, y1 + 1) for x in range(x0, x1 + 1)} if y1 >= y0 and x1 >= x0 else set()

    def sorted_diff(rn, ro):
        new = cells(rn) - cells(ro)
        if not new:
            return []
        y0, y1, x0, x1 = rn
        lists = []
        full_cols = [x for x in range(x0, x1 + 1) if all((y, x) in new for y in range(y0, y1 + 1))]
        for x in full_cols:
            lists.append(strip("c", x, y0, y1))
        rest = {c for c in new if c[1] not in full_cols}
        for y in sorted({y for y, _ in rest}):
            xs = sorted(x for yy, x in rest if yy == y)
            assert xs == list(range(xs[0], xs[-1] + 1))
            lists.append(strip("r", y, xs[0], xs[-1]))
        while len(lists) > 1:
            lists.sort(key=len)
            lists.append(merge(net, lists.pop(0), lists.pop(0)))
        return lists[0]

    outputs = {}

    def solve(group, r, L, t):
        m = N - len(cells(r))
        lo, hi = max(0, t - m), min(len(L) - 1, t)
        L, t = L[lo:hi + 1], t - lo
        if len(group) == 1:
            o = group[0]
            E = sorted_diff((o[0], o[0] + K - 1, o[1], o[1] + K - 1), r)
            outputs[o] = (merge(net, L, E) if L else E)[t] if E else L[t]
            return
        ys = sorted({o[0] for o in group})
        xs = sorted({o[1] for o in group})
        if len(xs) >= len(ys) and len(xs) > 1:
            mid = xs[len(xs) // 2]
            parts = [[o for o in group if o[1] < mid], [o for o in group if o[1] >= mid]]
        else:
            mid = ys[len(ys) // 2]
            parts = [[o for o in group if o[0] < mid], [o for o in group if o[0] >= mid]]
        for g in parts:
            rg = rect(g)
            E = sorted_diff(rg, r)
            solve(g, rg, (merge(net, L, E) if L else E) if E else L, t)

    outs = [(y, x) for y in range(TH) for x in range(TW)]
    r0 = rect(outs)
    solve(outs, r0, sorted_diff(r0, (0, -1, 0, -1)), T)
    return net, outputs


def live_ops(net, outputs):
    live, stack = set(), list(outputs.values())
    while stack:
        w = stack.pop()
        if w[0] == "op" and w[1] not in live:
            live.add(w[1])
            _, a, b = net.ops[w[1]]
            stack += [a, b]
    return sorted(live)


def evaluate(net, live, outputs, f):
    """f: (FH, FW, n) int array; returns {output: values}."""
    val = {}

    def get(w):
        return f[w[1][0], w[1][1]] if w[0] == "in" else val[w[1]]

    for i in live:
        kind, a, b = net.ops[i]
        val[i] = np.minimum(get(a), get(b)) if kind == "min" else np.maximum(get(a), get(b))
    return {o: get(w) for o, w in outputs.items()}


def check(net, live, outputs):
    rng = np.random.default_rng(5)
    for f in (rng.integers(0, 256, (FH, FW, 20000)), rng.integers(0, 3, (FH, FW, 20000)),
              rng.integers(0, 2, (FH, FW, 20000))):
        got = evaluate(net, live, outputs, f)
        for (y, x), v in got.items():
            want = np.median(f[y:y + K, x:x + K].reshape(N, -1), axis=0)
            assert np.array_equal(v, want), (y, x)


def emit(net, live, outputs, path):
    name = {}
    lines = []
    for n, i in enumerate(live):
        kind, a, b = net.ops[i]

        def ref(w):
            return "f[%d][%d]" % w[1] if w[0] == "in" else name[w[1]]

        name[i] = "t%d" % n
        lines.append("  const us2 t%d = __builtin_elementwise_%s(%s, %s);" % (
            n, kind, ref(a), ref(b)))
    outl = ["  out[%d][%d] = %s;" % (y, x, name[w[1]] if w[0] == "op" else "f[%d][%d]" % w[1])
            for (y, x), w in sorted(outputs.items())]
    text = """// Generated by gen_median5_net.py (in this directory); do not edit.
//
// Medians of the eight 5x5 windows of a 2-column x 4-row output tile:
// f[r][c] is footprint row r (0..7), column c (0..5); out[y][x] is the
// median of the window whose top-left footprint cell is (y, x).  %d packed
// min/max operations (16-bit lanes, two tiles side by side), checked
// against numpy's median by the generator.
#ifndef KCV_MEDIAN5_NET_H
#define KCV_MEDIAN5_NET_H

namespace kcv {
namespace filt {

__device__ __forceinline__ void median5_net_2x4(const us2 (&f)[8][6], us2 (&out)[4][2]) {
%s
%s
}

}  // namespace filt
}  // namespace kcv

#endif  // KCV_MEDIAN5_NET_H
""" % (len(live), "\n".join(lines), "\n".join(outl))
    with open(path, "w") as fh:
        fh.write(text)


def main():
    net, outputs = build()
    live = live_ops(net, outputs)
    check(net, live, outputs)
    here = os.path.dirname(os.path.abspath(__file__))
    emit(net, live, outputs, os.path.join(here, "median5_net.h"))
    print("median5_net.h: %d min/max operations for %d outputs" % (len(live), len(outputs)),
          file=sys.stderr)


if __name__ == "__main__":
    main()
