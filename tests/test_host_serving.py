"""Host logic of the request loop: the HBM arena and the request stream."""

import numpy as np

from lineax_amd import serving as S


def test_arena_first_fit_and_coalescing():
    ar = S.Arena.__new__(S.Arena)  # host-only: no device buffer
    ar.size, ar.free, ar.used, ar.high_water = 1 << 20, [(0, 1 << 20)], 0, 0
    rng = np.random.default_rng(0)
    live = {}
    for step in range(2000):
        if live and (rng.random() < 0.45 or len(live) > 40):
            key = list(live)[int(rng.integers(len(live)))]
            off, n = live.pop(key)
            ar.release(off, n)
        else:
            n = int(rng.integers(1, 40000))
            off = ar.alloc(n)
            if off is not None:
                live[step] = (off, n)
        # invariants: live blocks disjoint, free list sorted and coalesced, accounting exact
        spans = sorted((o, o + max(S.Arena.ALIGN, -(-n // S.Arena.ALIGN) * S.Arena.ALIGN)) for o, n in live.values())
        for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
            assert a1 <= b0
        for (o, ln), (o2, _) in zip(ar.free, ar.free[1:]):
            assert o + ln < o2
        assert ar.used == sum(e - s for s, e in spans)
        assert ar.used + sum(ln for _, ln in ar.free) == ar.size
    for off, n in list(live.values()):
        ar.release(off, n)
    assert ar.free == [(0, ar.size)] and ar.used == 0


def test_request_stream_statistics():
    reqs, ops = S.make_request_stream(2000.0, 400, seed=5, device="cpu", qr_m=(64, 96), chol_n=(16, 48))
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0) and arr[0] == 0.0
    gaps = np.diff(arr)
    assert abs(gaps.mean() * 2000.0 - 1.0) < 0.15
    kinds = [r.kind for r in reqs]
    assert 0.4 < kinds.count("qr") / len(kinds) < 0.6
    reuse = [r for r in reqs if r.reuse]
    assert 0.05 < len(reuse) / len(reqs) < 0.16
    for r in reqs:
        if r.kind == "qr":
            assert 64 <= r.m <= 96 and -(-r.m // 4) <= r.n <= r.m
        else:
            assert 16 <= r.n <= 48 and r.m == r.n
        a = r.a if r.a is not None else ops[r.op_id]
        assert tuple(a.shape) == (r.m, r.n) and tuple(r.b.shape) == (r.m,)
        if r.reuse:
            assert r.a is None and r.op_id in ops and r.op_id < r.rid


def test_request_flops_formulas():
    r = S.Request(rid=0, kind="qr", arrival=0.0, m=100, n=40, op_id=0, reuse=False, a=None, b=None)
    assert S.request_flops(r) == 2 * 100 * 40 * 40 - 2 * 40 ** 3 / 3 + 4 * 100 * 40
    r.reuse = True
    assert S.request_flops(r) == 4 * 100 * 40
    c = S.Request(rid=1, kind="chol", arrival=0.0, m=50, n=50, op_id=1, reuse=False, a=None, b=None)
    assert S.request_flops(c) == 50 ** 3 / 3 + 2 * 50 * 50
