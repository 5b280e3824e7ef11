"""Pins the oracle's streaming CPD pieces (CPU only).

  or_mttkrp_stream   equals the long-double COO MTTKRP (the reference's own
                     test oracle) and a dense einsum.
  or_kruskal_norm /  ||Z||^2 and the relative error against the dense model
  or_cpd_error       on small tensors (squared residual from numpy).
  or_stream_prepare  numpy lexsort with the stream mode most significant and
                     the others cyclically after it; labels are ranks of first
                     appearance in that order; iperm inverts them.
  or_stream_cpd      (a) the first slice's first time row equals the closed
                     form computed from the start factors, including the
                     reference's Gram of the *unnormalised* start factors
                     (grow_mats computes aTa before batch 0 is normalised);
                     (b) max-column-norm constraints hold on the output;
                     (c) on an exactly rank-2 stream every per-slice error
                     and final-err stay at a few percent (an unfitted model
                     sits near 1); (d) final-err equals the
                     error of the returned model on the whole stream;
                     (e) checkpoint errors (all slices so far, the last 10)
                     equal the dense squared error of that window against
                     the model after the slice (truncated-stream runs give
                     that model); (f) the refit equals CPD-ALS with the
                     default options from rand() values drawn right after
                     the slice, and consumes exactly one draw pair per
                     factor entry of the window.
"""
import ctypes

import numpy as np
import pytest

import oracle_py as orc

libc = ctypes.CDLL("libc.so.6")


def _rand_coo(dims, nnz, seed):
    rng = np.random.default_rng(seed)
    lin = rng.choice(int(np.prod(dims)), size=nnz, replace=False)
    inds = list(np.unravel_index(lin, dims))
    vals = rng.random(nnz) + 0.1
    return [np.asarray(i, dtype=np.uint64) for i in inds], vals


def _dense(inds, vals, dims):
    X = np.zeros(dims)
    for n in range(len(vals)):
        X[tuple(int(i[n]) for i in inds)] += vals[n]
    return X


def test_mttkrp_stream_matches_coo_and_dense():
    dims = [7, 5, 6]
    inds, vals = _rand_coo(dims, 80, 1)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    rng = np.random.default_rng(2)
    mats = [rng.standard_normal((d, 4)) for d in dims]
    X = _dense(inds, vals, dims)
    for mode in range(3):
        got = orc.mttkrp_stream(coo, mats, mode)
        ref = orc.mttkrp_coo(coo, mats, mode)
        others = [k for k in range(3) if k != mode]
        sub = "ijk,%sr,%sr->%sr" % ("ijk"[others[0]], "ijk"[others[1]], "ijk"[mode])
        dense = np.einsum(sub, X, mats[others[0]], mats[others[1]])
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-13)
        assert np.allclose(got, dense, rtol=1e-12, atol=1e-12)


def test_cpd_error_and_kruskal_norm():
    dims = [6, 5, 4]
    inds, vals = _rand_coo(dims, 50, 3)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    rng = np.random.default_rng(4)
    facs = [rng.standard_normal((d, 3)) for d in dims]
    lam = np.array([1.5, 0.5, 2.0])
    Z = np.einsum("r,ir,jr,kr->ijk", lam, *facs)
    X = _dense(inds, vals, dims)
    want = np.sqrt(np.sum((X - Z) ** 2) / np.sum(X ** 2))
    assert abs(orc.cpd_error(coo, facs, lam) - want) <= 1e-12 * max(1.0, want)


@pytest.mark.parametrize("stream_mode", [0, 1, 2])
def test_stream_prepare_order_and_labels(stream_mode):
    dims = [9, 7, 8]
    inds, vals = _rand_coo(dims, 150, 5 + stream_mode)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    iperm = orc.stream_prepare(coo, stream_mode)
    order = [(stream_mode + k) % 3 for k in range(3)]
    # expected: lexsort (last key most significant)
    keys = tuple(inds[m] for m in reversed(order))
    idx = np.lexsort(keys)
    got_inds = coo.inds()
    assert np.array_equal(coo.vals(), vals[idx])
    assert np.array_equal(got_inds[stream_mode], inds[stream_mode][idx])
    for m in range(3):
        if m == stream_mode:
            continue
        seq = inds[m][idx]
        _, first = np.unique(seq, return_index=True)
        appear = seq[np.sort(first)]
        label = {int(v): k for k, v in enumerate(appear)}
        assert np.array_equal(got_inds[m], np.array([label[int(v)] for v in seq], dtype=np.uint64))
        assert np.array_equal(iperm[m][:len(appear)], appear)


def _stream_tensor(T, d0, d1, seed, rank2=False):
    """Mode 2 is time with T slices; every slice non-empty."""
    rng = np.random.default_rng(seed)
    if rank2:
        A = rng.random((d0, 2))
        B = rng.random((d1, 2))
        C_ = rng.random((T, 2)) + 0.5
        X = np.einsum("ir,jr,kr->ijk", A, B, C_)
        ii, jj, kk = np.nonzero(X > -1)
        inds = [ii.astype(np.uint64), jj.astype(np.uint64), kk.astype(np.uint64)]
        return inds, X[ii, jj, kk], [d0, d1, T]
    inds, vals = [], []
    cells = []
    for t in range(T):
        n = int(rng.integers(5, 20))
        lin = rng.choice(d0 * d1, size=n, replace=False)
        cells.append((np.full(n, t), lin // d1, lin % d1, rng.random(n) + 0.1))
    kk = np.concatenate([c[0] for c in cells]).astype(np.uint64)
    ii = np.concatenate([c[1] for c in cells]).astype(np.uint64)
    jj = np.concatenate([c[2] for c in cells]).astype(np.uint64)
    vv = np.concatenate([c[3] for c in cells])
    return [ii, jj, kk], vv, [d0, d1, T]


CONS = [(2, 1, 0.0), (2, 1, 0.0), (1, 0, 1e-2)]   # maxcolnorm x2, frob on time


def test_stream_first_time_row_closed_form():
    inds, vals, dims = _stream_tensor(1, 6, 5, 7)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    orc.stream_prepare(coo, 2)
    R = 3
    libc.srand(11)
    res = orc.stream_cpd(coo, 2, R, 0.95, 1, CONS)
    # the start factors: rand() rows for mode 0 then mode 1 (batch dims)
    libc.srand(11)
    d0 = int(max(coo.inds()[0]) + 1)
    d1 = int(max(coo.inds()[1]) + 1)
    U0 = orc.fill_rand(d0 * R).reshape(d0, R)
    U1 = orc.fill_rand(d1 * R).reshape(d1, R)
    G = (U0.T @ U0) * (U1.T @ U1) + 1e-2 * np.eye(R)        # unnormalised Grams
    N0 = U0 / np.linalg.norm(U0, axis=0)
    N1 = U1 / np.linalg.norm(U1, axis=0)
    ii, jj = coo.inds()[0], coo.inds()[1]
    m = np.sum(coo.vals()[:, None] * N0[ii.astype(int)] * N1[jj.astype(int)], axis=0)
    want = np.linalg.solve(G, m)
    assert np.allclose(res["time"][0], want, rtol=1e-10, atol=1e-12)


def test_stream_constraints_and_final_error():
    inds, vals, dims = _stream_tensor(12, 10, 8, 9)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    orc.stream_prepare(coo, 2)
    libc.srand(3)
    res = orc.stream_cpd(coo, 2, 4, 0.95, 5, CONS)
    assert res["nbatches"] == 12
    for m in (0, 1):
        assert np.all(np.linalg.norm(res["factors"][m], axis=0) <= 1 + 1e-9)
    assert np.all(res["outer_its"] >= 1) and np.all(res["outer_its"] <= 5)
    facs = [res["factors"][0], res["factors"][1], res["time"]]
    e = orc.cpd_error(coo, facs, np.ones(4))
    assert abs(res["final_err"] - e * e) <= 1e-12
    assert np.all(res["local_err"] >= 0)


def test_stream_fits_exact_rank2():
    inds, vals, dims = _stream_tensor(40, 8, 6, 13, rank2=True)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    orc.stream_prepare(coo, 2)
    libc.srand(5)
    res = orc.stream_cpd(coo, 2, 2, 0.95, 10, [(2, 1, 0.0), (2, 1, 0.0), (1, 0, 1e-6)])
    # the stream's loose tolerances (outer 5e-2, ADMM 7e-2) leave a few
    # percent; a model that had not learned the structure sits near 1
    assert np.max(res["local_err"]) < 0.06
    assert res["final_err"] < 0.05


def _truncate(coo, k, t=2):
    """The prepared stream's first k slices as a stream of its own (labels
    kept: they are first-appearance labels already)."""
    inds = coo.inds()
    keep = inds[t] < k
    sub = [i[keep] for i in inds]
    dims = [int(i.max()) + 1 for i in sub]
    dims[t] = k
    return orc.Coo.from_arrays(sub, coo.vals()[keep], dims), sub, coo.vals()[keep], dims


def test_stream_checkpoint_errors_dense():
    """global / local-10 errors at checkpoints equal the dense squared error
    of the window against the model after that slice."""
    inds, vals, dims = _stream_tensor(25, 10, 8, 17)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    orc.stream_prepare(coo, 2)
    R = 3
    libc.srand(3)
    full = orc.stream_cpd(coo, 2, R, 0.95, 5, CONS, check_interval=12)
    at = {11, 23, 24}                       # slices 12, 24 and the last one
    for b in range(25):
        for key in ("global_err", "local10_err"):
            assert (full[key][b] >= 0) == (b in at), (key, b)
        assert full["cpd_err"][b] == -1.0
    assert abs(full["global_err"][24] - full["final_err"]) <= 1e-13
    for k in (12, 24, 25):
        sub_coo, sub, sv, sd = _truncate(coo, k)
        libc.srand(3)
        tr = orc.stream_cpd(sub_coo, 2, R, 0.95, 5, CONS)
        assert tr["nbatches"] == k
        A, B, Tm = tr["factors"][0], tr["factors"][1], tr["time"]
        for b0, key in ((0, "global_err"), (max(0, k - 10), "local10_err")):
            w = sub[2] >= b0
            X = _dense([sub[0][w], sub[1][w], sub[2][w] - b0], sv[w], [sd[0], sd[1], k - b0])
            Z = np.einsum("ir,jr,kr->ijk", A, B, Tm[b0:k])
            want = np.sum((X - Z) ** 2) / np.sum(X ** 2)
            assert abs(full[key][k - 1] - want) <= 1e-10 * want, (k, key)


def test_stream_refit_window_options_and_draws():
    """The refit at a checkpoint is CPD-ALS (tolerance 1e-5, 200 iterations)
    of the slices so far, started from rand() values drawn right after that
    slice for every mode in order; later rows are drawn after those."""
    inds, vals, dims = _stream_tensor(15, 9, 7, 21)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    orc.stream_prepare(coo, 2)
    R = 3
    libc.srand(7)
    full = orc.stream_cpd(coo, 2, R, 0.95, 5, CONS, check_interval=6, refit=True)
    after_full = libc.rand()
    assert [b for b in range(15) if full["cpd_err"][b] != -1.0] == [5, 11, 14]
    # the refit at slice 6: the rand() state is that after slices 0..5
    sub_coo, sub, sv, sd = _truncate(coo, 6)
    libc.srand(7)
    orc.stream_cpd(sub_coo, 2, R, 0.95, 5, CONS)
    start = [orc.fill_rand(d * R).reshape(d, R) for d in sd]
    csf = orc.Csf(sub_coo, orc.mode_order(sd, 0))
    _, _, fits = orc.cpd_als([csf], [0, 0, 0], start, 200, 1e-5)
    assert 0 < len(fits) <= 200
    assert abs(full["cpd_err"][5] - (1.0 - fits[-1])) <= 1e-12
    assert 0.0 < full["cpd_err"][5] < 1.0
    # draws: the same stream without refits, then one refit's worth of
    # draws per checkpoint, leaves rand() at the same place
    libc.srand(7)
    plain = orc.stream_cpd(coo, 2, R, 0.95, 5, CONS, check_interval=6)
    assert np.array_equal(plain["local_err"][:6], full["local_err"][:6])
    for k in (6, 12, 15):
        _, _, _, kd = _truncate(coo, k)
        for _ in range(2 * R * sum(kd)):
            libc.rand()
    assert libc.rand() == after_full
