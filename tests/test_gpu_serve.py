"""The multi-stream server (no reference counterpart).

  alone == offline   one stream through the server equals the whole-stream
                     entry (splatt_amd_cpd_stream_coo, itself checked against
                     the oracle) on the same slices and seed.
  mixed == alone     several streams of different shapes and ranks submitted
                     interleaved, with several slices in flight, give each
                     stream the result it gets alone (1e-9 relative: only the
                     order of floating-point atomics in the COO MTTKRP may
                     differ between runs).
  admission          with a budget that fits two streams, the others wait and
                     are admitted as streams close; every slice completes and
                     the accounted memory never exceeds the budget.
"""
import numpy as np
import pytest

import pysplatt as ps

pytestmark = pytest.mark.gpu


def _slices(T, dims, per, seed, zipf=None):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(T):
        n = int(rng.integers(per[0], per[1] + 1))
        cols = []
        for d in dims:
            if zipf:
                cols.append(np.minimum(rng.zipf(zipf, size=4 * n) - 1, d - 1))
            else:
                cols.append(rng.integers(0, d, size=4 * n))
        key = np.zeros(4 * n, dtype=np.int64)
        for c, d in zip(cols, dims):
            key = key * d + c
        _, first = np.unique(key, return_index=True)
        first = rng.permutation(np.sort(first)[:n])
        inds = [c[first].astype(np.uint64) for c in cols] + [np.full(len(first), t, dtype=np.uint64)]
        out.append((inds, rng.random(len(first)) + 0.05))
    return out


def _serve(streams, max_inflight, interleave=True, budget=0):
    sv = ps.Server(budget=budget, max_inflight=max_inflight)
    ids = [sv.open(s["cap"] + [1], s["R"], 2, forget=0.95, max_its=5, frob=1e-2, seed=s["seed"])
           for s in streams]
    tickets = []
    T = max(len(s["slices"]) for s in streams)
    if interleave:
        for t in range(T):
            for sid, s in zip(ids, streams):
                if t < len(s["slices"]):
                    tickets.append(sv.submit(sid, *s["slices"][t]))
    else:
        for sid, s in zip(ids, streams):
            for sl in s["slices"]:
                tickets.append(sv.submit(sid, *sl))
    for sid in ids:
        sv.close(sid)
    sv.drain()
    info = [sv.slice(t) for t in tickets]
    res = []
    for sid, s in zip(ids, streams):
        res.append([sv.factors(sid, m) for m in range(3)])
    stats = sv.stats()
    sv.free()
    return res, info, stats


def _streams():
    specs = [([300, 200], 8, (200, 600)), ([1000, 50], 16, (300, 900)),
             ([80, 700], 4, (100, 300)), ([500, 500], 32, (400, 800)),
             ([2000, 1500], 8, (500, 1500)), ([120, 90], 24, (50, 200)),
             ([900, 700], 64, (600, 1200))]
    out = []
    for k, (cap, R, per) in enumerate(specs):
        out.append(dict(cap=cap, R=R, seed=100 + k,
                        slices=_slices(4 + k % 3, cap, per, seed=k, zipf=1.3 if k % 2 else None)))
    return out


def test_single_stream_equals_offline(gpu_available):
    s = _streams()[1]
    res, info, _ = _serve([s], 1)
    inds = [np.concatenate([sl[0][m] for sl in s["slices"]]) for m in range(3)]
    vals = np.concatenate([sl[1] for sl in s["slices"]])
    dims = [s["cap"][0], s["cap"][1], len(s["slices"])]
    off = ps.run_stream(inds, vals, dims, s["R"], 2, forget=0.95, max_its=5, seed=s["seed"],
                        frob=1e-2)
    assert off is not None
    for m in range(3):
        want = off["factors"][m]
        got = res[0][m]
        assert got.shape == want.shape
        assert np.allclose(got, want, rtol=1e-9, atol=1e-12 * max(1.0, np.max(np.abs(want))))
    assert np.allclose([i["local_err"] for i in info], off["local_err"], rtol=1e-9, atol=1e-12)


def test_mixed_equals_alone(gpu_available):
    streams = _streams()
    mixed, info, _ = _serve(streams, 4, interleave=True)
    assert all(i is not None for i in info)
    for k, s in enumerate(streams):
        alone, _, _ = _serve([s], 1)
        for m in range(3):
            a, b = alone[0][m], mixed[k][m]
            assert a.shape == b.shape
            assert np.allclose(a, b, rtol=1e-9, atol=1e-12 * max(1.0, np.max(np.abs(a)))), (k, m)


def test_admission_by_budget(gpu_available):
    streams = _streams()[:4]
    # the footprint of the two largest streams
    sv = ps.Server(budget=0, max_inflight=2)
    foot = []
    for s in streams:
        sid = sv.open(s["cap"] + [1], s["R"], 2, seed=1)
        foot.append(sv.stats()["used"])      # the only open stream
        sv.close(sid)
    sv.free()
    budget = sorted(foot)[-1] + sorted(foot)[-2]
    res, info, stats = _serve(streams, 2, interleave=False, budget=budget)
    assert all(i is not None for i in info)
    assert stats["used"] == 0 and stats["waiting"] == 0
    alone = [_serve([s], 1)[0][0] for s in streams]
    for k in range(4):
        for m in range(3):
            assert np.allclose(res[k][m], alone[k][m], rtol=1e-9,
                               atol=1e-12 * max(1.0, np.max(np.abs(alone[k][m]))))


def test_served_rank32_stream_matches_oracle(gpu_available):
    # a rank-32 Zipf(1.2) stream served alone against the oracle's streaming
    # CPD: the server draws new rows from the stream's own rand() sequence,
    # which is glibc's after srand(seed)
    import ctypes
    import oracle_py as orc
    libc = ctypes.CDLL("libc.so.6")
    cap = [3000, 2000]
    s = dict(cap=cap, R=32, seed=77, slices=_slices(6, cap, (1500, 3000), seed=77, zipf=1.2))
    res, info, _ = _serve([s], 1)
    inds = [np.concatenate([sl[0][m] for sl in s["slices"]]) for m in range(3)]
    vals = np.concatenate([sl[1] for sl in s["slices"]])
    dims = [int(inds[0].max()) + 1, int(inds[1].max()) + 1, len(s["slices"])]
    coo = orc.Coo.from_arrays(inds, vals, dims)
    orc.stream_prepare(coo, 2)
    libc.srand(77)
    ref = orc.stream_cpd(coo, 2, 32, 0.95, 5, [(2, 1, 0.0), (2, 1, 0.0), (1, 0, 1e-2)])
    assert [i["outer"] for i in info] == [int(x) for x in ref["outer_its"]]
    assert np.allclose([i["local_err"] for i in info], ref["local_err"], rtol=1e-8, atol=1e-12)
    for m, want in ((0, ref["factors"][0]), (1, ref["factors"][1]), (2, ref["time"])):
        got = res[0][m]
        assert got.shape == want.shape, m
        assert np.allclose(got, want, rtol=1e-8, atol=1e-8 * np.max(np.abs(want))), m
