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
                     factor entry of the window; (g) with closed-form
                     (frob) modes, the time row and each mode's update at a
                     later slice make the gradient of the forgetting-weighted
                     streaming objective vanish, the objective built from
                     dense reconstructions of the slices (pins the
                     forgetting factor and the orientation of the history
                     cross-Gram); (h) with two outer iterations, the history
                     holds every outer iteration's time row (pins the time
                     Gram update inside the outer loop); (i) max-column-norm
                     mode updates equal the ADMM solve, from a zero dual, of
                     the normal equations read off that objective's
                     gradient (pins the dual reset before each mode).
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


# ---- the per-slice update against the streaming objective -----------------
#
# StreamCPD.cc:406-492 updates, per slice t, the time row and then every
# other mode. Each update minimises a quadratic whose terms are fixed by the
# streaming model (Smith et al., SDM'18): the slice's own squared residual,
# the history term sum_g ||[[A_old, B_old, g]] - [[A, B, g]]||^2 over
# weighted time rows g (the forgetting factor f^(age) on every earlier
# slice, and -- the time Gram being updated inside the outer loop -- the
# current slice's earlier outer-iteration rows with weight 1), and the frob
# penalty. The checks below build that objective from dense reconstructions
# of the slices (numpy, no Gram algebra), take its gradient, and require it
# to vanish at the oracle's output (closed-form modes), or derive its normal
# equations from the gradient and require the oracle's mode update to be the
# ADMM solve of them from a zero dual (max-column-norm modes).

def _stream_full_first(T, d0, d1, seed, big=1.0):
    """Time mode 2 with T slices; slice 0 touches every row of modes 0 and 1,
    so the factors never grow after it."""
    rng = np.random.default_rng(seed)
    ii, jj, kk, vv = [], [], [], []
    first = set()
    for i in range(d0):
        first.add((i, int(rng.integers(d1))))
    for j in range(d1):
        first.add((int(rng.integers(d0)), j))
    for t in range(T):
        cells = set(first) if t == 0 else set()
        n = int(rng.integers(d0 * d1 // 4, d0 * d1 // 2))
        for lin in rng.choice(d0 * d1, size=n, replace=False):
            cells.add((int(lin) // d1, int(lin) % d1))
        for (i, j) in sorted(cells):
            ii.append(i)
            jj.append(j)
            kk.append(t)
            vv.append((rng.random() + 0.1) * big)
    inds = [np.array(ii, dtype=np.uint64), np.array(jj, dtype=np.uint64),
            np.array(kk, dtype=np.uint64)]
    return inds, np.array(vv), [d0, d1, T]


def _slice_dense(sub, sv, sd, s):
    w = sub[2] == s
    X = np.zeros((sd[0], sd[1]))
    np.add.at(X, (sub[0][w].astype(int), sub[1][w].astype(int)), sv[w])
    return X


def _recon(A, B, c):
    return (A * c) @ B.T


def _grad(which, A, B, c, X, hist, Aold, Bold, lam):
    """Gradient in A (which=0) or B (which=1) of
    1/2 ||X - [[A,B,c]]||^2 + 1/2 sum_g ||[[Aold,Bold,g]] - [[A,B,g]]||^2
    + lam/2 ||A or B||^2."""
    terms = [(_recon(A, B, c) - X, c)]
    terms += [(_recon(A, B, g) - _recon(Aold, Bold, g), g) for g in hist]
    if which == 0:
        return sum(E @ (B * g) for E, g in terms) + lam * A
    return sum(E.T @ (A * g) for E, g in terms) + lam * B


def _normal_equations(which, A, B, c, X, hist, Aold, Bold, lam):
    """(G, M) with gradient(U) = U G - M, read off the gradient of the
    objective at zero and at unit rows (the Hessian is the same for every
    row)."""
    R = A.shape[1]
    Z = np.zeros_like(A if which == 0 else B)
    args = lambda U: (U, B) if which == 0 else (A, U)
    M = -_grad(which, *args(Z), c, X, hist, Aold, Bold, lam)
    G = np.zeros((R, R))
    for r in range(R):
        E = Z.copy()
        E[0, r] = 1.0
        G[r] = _grad(which, *args(E), c, X, hist, Aold, Bold, lam)[0] + M[0]
    return G, M


def _runs(coo, T, seed, max_its, cons, forget=0.95):
    """The oracle run on the first k slices, for k = 1 .. T."""
    out = {}
    for k in range(1, T + 1):
        sub_coo, sub, sv, sd = _truncate(coo, k)
        libc.srand(seed)
        out[k] = orc.stream_cpd(sub_coo, 2, 3, forget, max_its, cons)
    return out, sub, sv, sd


def _time_row(A, B, X, lam_t, gram_from=None):
    """argmin_c 1/2 ||X - [[A,B,c]]||^2 + lam_t/2 ||c||^2 by least squares
    on the unfolded slice. gram_from replaces the design's Gram (the first
    slice's quirk: its Gram comes from the unnormalised start factors)."""
    K = np.einsum("ir,jr->ijr", A, B).reshape(-1, A.shape[1])
    KtK = K.T @ K if gram_from is None else gram_from
    return np.linalg.solve(KtK + lam_t * np.eye(A.shape[1]), K.T @ X.reshape(-1))


FROB = [(1, 0, 1e-4), (1, 0, 2e-4), (1, 0, 1e-2)]


def test_stream_update_is_stationary_for_the_forgetting_objective():
    """One outer iteration per slice, frob (closed form) on every mode: at
    the last slice the time row, then A, then B are exact minimisers of the
    slice + forgetting-weighted history objective. A dropped forgetting
    factor or a transposed history cross-Gram breaks it."""
    T, f = 4, 0.95
    inds, vals, dims = _stream_full_first(T, 7, 6, 31)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    orc.stream_prepare(coo, 2)
    runs, sub, sv, sd = _runs(coo, T, 9, 1, FROB, f)
    prev, last = runs[T - 1], runs[T]
    assert np.all(last["outer_its"] == 1)
    Ap, Bp = prev["factors"][:2]
    A, B = last["factors"][:2]
    assert Ap.shape == A.shape and Bp.shape == B.shape    # no growth
    c = last["time"]
    X = _slice_dense(sub, sv, sd, T - 1)
    hist = [np.sqrt(f ** (T - 1 - s)) * c[s] for s in range(T - 1)]
    # time row from the factors at the slice's start
    ct = _time_row(Ap, Bp, X, FROB[2][2])
    assert np.allclose(c[T - 1], ct, rtol=1e-10, atol=1e-12)
    gA = _grad(0, A, Bp, c[T - 1], X, hist, Ap, Bp, FROB[0][2])
    gB = _grad(1, A, B, c[T - 1], X, hist, Ap, Bp, FROB[1][2])
    scale = np.max(np.abs(X)) * np.max(np.abs(c))
    assert np.max(np.abs(gA)) <= 1e-10 * scale
    assert np.max(np.abs(gB)) <= 1e-10 * scale
    # the same objective without forgetting, or with the history cross-Gram
    # transposed, is not stationary at the output
    flat = [c[s] for s in range(T - 1)]
    assert np.max(np.abs(_grad(1, A, B, c[T - 1], X, flat, Ap, Bp, FROB[1][2]))) > 1e-4 * scale
    Gt, Mt = _normal_equations(1, A, B, c[T - 1], X, hist, Ap, Bp, FROB[1][2])
    assert np.allclose(B @ Gt, Mt, rtol=1e-9, atol=1e-10 * scale)


def test_stream_time_gram_takes_every_outer_iteration():
    """Two outer iterations per slice: the history of slice t holds both
    time rows of every earlier slice (each with its forgetting weight) and
    the first-iteration row of slice t itself (StreamCPD.cc:420-435 adds
    each outer iteration's row to the time Gram). The first-iteration rows
    are not returned; they are recomputed as least-squares time rows from
    the factors at each slice's start."""
    T, f, R = 3, 0.95, 3
    inds, vals, dims = _stream_full_first(T, 7, 6, 37)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    orc.stream_prepare(coo, 2)
    runs, sub, sv, sd = _runs(coo, T, 4, 2, FROB, f)
    assert np.all(runs[T]["outer_its"] == 2)
    lam_t = FROB[2][2]
    # start factors of slice 0: rand() rows, normalised, Gram unnormalised
    libc.srand(4)
    U0 = orc.fill_rand(sd[0] * R).reshape(sd[0], R)
    U1 = orc.fill_rand(sd[1] * R).reshape(sd[1], R)
    G0 = (U0.T @ U0) * (U1.T @ U1)
    N0 = U0 / np.linalg.norm(U0, axis=0)
    N1 = U1 / np.linalg.norm(U1, axis=0)
    first = [_time_row(N0, N1, _slice_dense(sub, sv, sd, 0), lam_t, gram_from=G0)]
    for s in range(1, T):
        As, Bs = runs[s]["factors"][:2]
        first.append(_time_row(As, Bs, _slice_dense(sub, sv, sd, s), lam_t))
    c = runs[T]["time"]
    hist = []
    for s in range(T - 1):
        w = np.sqrt(f ** (T - 1 - s))
        hist += [w * first[s], w * c[s]]
    hist.append(first[T - 1])
    Ap, Bp = runs[T - 1]["factors"][:2]
    A, B = runs[T]["factors"][:2]
    X = _slice_dense(sub, sv, sd, T - 1)
    scale = np.max(np.abs(X)) * np.max(np.abs(c))
    gB = _grad(1, A, B, c[T - 1], X, hist, Ap, Bp, FROB[1][2])
    assert np.max(np.abs(gB)) <= 1e-9 * scale
    # one row per slice (the final one) does not give a stationary point
    once = [np.sqrt(f ** (T - 1 - s)) * c[s] for s in range(T - 1)]
    assert np.max(np.abs(_grad(1, A, B, c[T - 1], X, once, Ap, Bp, FROB[1][2]))) > 1e-4 * scale


def test_stream_mode_update_is_admm_from_a_zero_dual():
    """Max-column-norm modes (the streaming default): each mode update at a
    slice is the ADMM solve (25 iterations at most, tolerance 7e-2, the
    stream's settings) of the objective's normal equations, started from
    the mode's factor and a zero dual (StreamCPD.cc:466-469 clears the dual
    before every mode's ADMM). The previous slice's ADMM ended with an active
    column constraint, so a dual carried over would change the result."""
    T, f = 4, 0.95
    inds, vals, dims = _stream_full_first(T, 7, 6, 41, big=3.0)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    orc.stream_prepare(coo, 2)
    runs, sub, sv, sd = _runs(coo, T, 12, 1, CONS, f)
    Ap, Bp = runs[T - 1]["factors"][:2]
    A, B = runs[T]["factors"][:2]
    c = runs[T]["time"]
    # the previous slice left a column on the constraint (norm 1): its
    # ADMM's final projection was active and its dual non-zero
    assert np.any(np.abs(np.linalg.norm(Ap, axis=0) - 1.0) < 1e-12)
    assert np.any(np.abs(np.linalg.norm(Bp, axis=0) - 1.0) < 1e-12)
    X = _slice_dense(sub, sv, sd, T - 1)
    hist = [np.sqrt(f ** (T - 1 - s)) * c[s] for s in range(T - 1)]
    GA, MA = _normal_equations(0, Ap, Bp, c[T - 1], X, hist, Ap, Bp, 0.0)
    wantA, _, _, _ = orc.admm(1, MA, Ap, np.zeros_like(Ap), GA, 25, 7e-2)
    assert np.allclose(A, wantA, rtol=1e-10, atol=1e-12)
    GB, MB = _normal_equations(1, A, Bp, c[T - 1], X, hist, Ap, Bp, 0.0)
    wantB, _, _, _ = orc.admm(1, MB, Bp, np.zeros_like(Bp), GB, 25, 7e-2)
    assert np.allclose(B, wantB, rtol=1e-10, atol=1e-12)


def test_stream_oracle_threads_do_not_change_results():
    # the oracle's row-parallel loops (triangular solves, history product)
    # give the same bits at any thread count
    import oracle_py as orc
    rng = np.random.default_rng(5)
    n = 4000
    inds = [rng.integers(0, 900, n).astype(np.uint64), rng.integers(0, 700, n).astype(np.uint64),
            rng.integers(0, 4, n).astype(np.uint64)]
    vals = rng.random(n)
    outs = []
    for thr in (1, 4):
        coo = orc.Coo.from_arrays(inds, vals, [900, 700, 4])
        orc.stream_prepare(coo, 2)
        orc.set_threads(thr)
        try:
            import ctypes
            ctypes.CDLL("libc.so.6").srand(3)
            outs.append(orc.stream_cpd(coo, 2, 8, 0.95, 5, [(2, 1, 0.0), (2, 1, 0.0), (1, 0, 1e-2)]))
        finally:
            orc.set_threads(1)
    for key in ("time", "local_err"):
        assert np.array_equal(outs[0][key], outs[1][key]), key
    for m in range(2):
        assert np.array_equal(outs[0]["factors"][m], outs[1]["factors"][m]), m
