"""Pins the oracle's dense kernels, the rand() initialisation and CPD-ALS.

  or_mat_aTa           numpy A.T @ A.
  or_cholesky/_solve   L L^T reconstructs G, L equals numpy's Cholesky, the
                       solve's residual ||G x - b|| is at rounding level.
  or_mat_normalize     unit column norms and A_out * lambda == A_in.
  or_fill_rand         the glibc rand() stream called directly: each value
                       consumes two draws, magnitude 3*r/RAND_MAX, negative
                       when the second draw is even (src/util.c:15-24).
  or_cpd_als           (a) one sweep against a textbook ALS written with
                       numpy (Khatri-Rao + normal equations) from the same
                       start; (b) the reported fit of the last iteration
                       equals 1 - ||X - X_hat||^2/||X||^2 computed on the
                       dense tensor; (c) fit non-decreasing across iterations
                       (the reference asserts err <= olderr,
                       src/cpd/cpd.c:258); (d) an exactly rank-2 tensor is
                       fitted to within 1e-6.
"""
import ctypes

import numpy as np
import pytest

import oracle_py as orc

libc = ctypes.CDLL("libc.so.6")
RAND_MAX = 2147483647


def test_aTa_and_normalize():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((37, 5))
    out = np.zeros((5, 5))
    orc.lib().or_mat_aTa(orc._p(A, ctypes.c_double), 37, 5, orc._p(out, ctypes.c_double))
    assert np.allclose(out, A.T @ A, rtol=1e-13, atol=1e-12)
    B = A.copy()
    lam = np.zeros(5)
    orc.lib().or_mat_normalize(orc._p(B, ctypes.c_double), 37, 5, orc._p(lam, ctypes.c_double))
    assert np.allclose(np.linalg.norm(B, axis=0), 1.0)
    assert np.allclose(B * lam, A)


def test_cholesky_and_solve():
    rng = np.random.default_rng(2)
    M = rng.standard_normal((9, 9))
    G = M @ M.T + 9 * np.eye(9)
    L = G.copy()
    assert orc.lib().or_cholesky(orc._p(L, ctypes.c_double), 9) == 0
    low = np.tril(L)
    assert np.allclose(low @ low.T, G, rtol=1e-12, atol=1e-12)
    assert np.allclose(low, np.linalg.cholesky(G), rtol=1e-12, atol=1e-12)
    Bm = rng.standard_normal((13, 9))
    X = Bm.copy()
    orc.lib().or_cholesky_solve(orc._p(L, ctypes.c_double), 9, orc._p(X, ctypes.c_double), 13)
    assert np.max(np.abs(X @ G - Bm)) <= 1e-10
    # a non-positive pivot is reported like LAPACK's info
    bad = -np.eye(3)
    assert orc.lib().or_cholesky(orc._p(bad, ctypes.c_double), 3) == 1


def test_fill_rand_stream():
    libc.srand(1234)
    got = orc.fill_rand(20)
    libc.srand(1234)
    for v in got:
        a = libc.rand()
        b = libc.rand()
        mag = 3.0 * a / RAND_MAX
        assert abs(abs(v) - mag) <= 1e-15
        assert (v < 0) == (b % 2 == 0) or mag == 0.0


def _dense(inds, vals, dims):
    X = np.zeros(dims)
    np.add.at(X, tuple(inds), vals)
    return X


def _khatri_rao(mats):
    R = mats[0].shape[1]
    out = mats[0]
    for M in mats[1:]:
        out = (out[:, None, :] * M[None, :, :]).reshape(-1, R)
    return out


def _numpy_als_sweep(X, facs):
    """One textbook ALS sweep with per-mode 2-norm column normalisation."""
    N = X.ndim
    facs = [f.copy() for f in facs]
    lam = None
    for m in range(N):
        others = [facs[k] for k in range(N) if k != m]
        Xm = np.moveaxis(X, m, 0).reshape(X.shape[m], -1)
        KR = _khatri_rao(others)
        G = np.ones((facs[0].shape[1],) * 2)
        for o in others:
            G *= o.T @ o
        A = np.linalg.solve(G, (Xm @ KR).T).T
        lam = np.linalg.norm(A, axis=0)
        facs[m] = A / lam
    return facs, lam


def _tensor(dims, nnz, seed, rank_r=None):
    rng = np.random.default_rng(seed)
    if rank_r is None:
        cells = rng.choice(int(np.prod(dims)), nnz, replace=False)
        inds = list(np.unravel_index(cells, dims))
        vals = rng.random(nnz) + 0.1
    else:
        fs = [rng.random((d, rank_r)) + 0.2 for d in dims]
        X = np.einsum("ir,jr,kr->ijk", *fs)
        inds = list(np.indices(dims).reshape(len(dims), -1))
        vals = X.reshape(-1)
    return inds, vals


def _init(dims, R, seed):
    libc.srand(seed)
    return [orc.fill_rand(d * R).reshape(d, R) for d in dims]


def test_cpd_one_sweep_matches_numpy():
    dims = [6, 5, 7]
    inds, vals = _tensor(dims, 90, 3)
    c = orc.Coo.from_arrays(inds, vals, dims)
    csfs = [orc.Csf(c, orc.mode_order(dims, 3, m)) for m in range(3)]
    init = _init(dims, 3, 42)
    facs, lam, fits = orc.cpd_als(csfs, [0, 1, 2], init, 1, 0.0)
    # the oracle normalises the start first (src/cpd/cpd.c:179-188)
    start = [f / np.linalg.norm(f, axis=0) for f in init]
    want, wl = _numpy_als_sweep(_dense(inds, vals, dims), start)
    # after the final post-processing the oracle's factors are unit columns
    # and lambda carries the norms: compare the models
    for m in range(3):
        assert np.allclose(facs[m], want[m], rtol=1e-9, atol=1e-9)
    assert np.allclose(lam, wl, rtol=1e-9)


@pytest.mark.parametrize("seed", [1, 2])
def test_cpd_fit_matches_dense_residual_and_is_monotone(seed):
    dims = [8, 6, 7]
    inds, vals = _tensor(dims, 150, seed)
    c = orc.Coo.from_arrays(inds, vals, dims)
    csfs = [orc.Csf(c, orc.mode_order(dims, 0)), orc.Csf(c, orc.mode_order(dims, 3, 1))]
    perm0 = orc.mode_order(dims, 0)
    mode_csf = [1 if perm0[-1] == m else 0 for m in range(3)]
    facs, lam, fits = orc.cpd_als(csfs, mode_csf, _init(dims, 4, seed), 15, 0.0)
    assert len(fits) == 15
    assert np.all(np.diff(fits) >= -1e-12)
    X = _dense(inds, vals, dims)
    Xh = np.einsum("r,ir,jr,kr->ijk", lam, *facs)
    fit = 1 - np.sum((X - Xh) ** 2) / np.sum(X ** 2)
    assert abs(fit - fits[-1]) <= 1e-10


def test_cpd_exact_low_rank():
    dims = [7, 6, 5]
    inds, vals = _tensor(dims, None, 11, rank_r=2)
    c = orc.Coo.from_arrays(inds, vals, dims)
    csfs = [orc.Csf(c, orc.mode_order(dims, 3, m)) for m in range(3)]
    facs, lam, fits = orc.cpd_als(csfs, [0, 1, 2], _init(dims, 2, 5), 200, 1e-12)
    assert fits[-1] >= 1 - 1e-6


# ---- constraints and ADMM (src/cpd/admm.c, src/cpd/constraints/*.c) ------
#
#   or_admm, prox none   with the identity prox the dual vanishes after one
#                        step and P <- (G + rho I)^{-1}(M + rho P) is a
#                        fixed-point iteration whose limit is G^{-1} M.
#   or_admm, nonneg      the limit satisfies the KKT conditions of the
#                        non-negative least squares problem per row:
#                        p >= 0, g = G p - m >= 0 where p = 0, g = 0 where
#                        p > 0 (a wrong dual sign or prox placement fails).
#   or_admm, max-col-norm  when the unconstrained solution has column norms
#                        below 1 the constraint is inactive: the limit is
#                        G^{-1} M; otherwise every column norm is <= 1.
#   or_prox              closed forms of the three operators.
#   or_admm_chunked      equals or_admm run on each chunk of rows on its own
#                        (ragged last chunk, summed iteration count); one
#                        chunk of at least I rows is the one-block ADMM; a
#                        chunked nonneg CPD meets the nonneg least-squares
#                        optimality conditions.
#   or_cpd_als_con       frob with multiplier 0 equals the unconstrained
#                        solve without normalisation; with multiplier mu the
#                        solved factor satisfies (G + mu I) U^T = M^T.


def _spd(R, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((3 * R, R))
    return scale * (A.T @ A) + 0.1 * np.eye(R)


def test_prox_closed_forms():
    rng = np.random.default_rng(4)
    P = rng.standard_normal((50, 6)) * 0.5
    P[:, 0] *= 10.0
    nn = orc.prox(3, P)
    assert np.array_equal(nn, np.where(P > 0, P, 0.0))
    mc = orc.prox(1, P)
    norms = np.linalg.norm(P, axis=0)
    want = np.where(norms > 1, P / np.where(norms > 1, norms, 1.0), P)
    assert np.allclose(mc, want, rtol=1e-15, atol=0)
    mcn = orc.prox(2, P)
    Q = np.where(P < 0, 0.0, P)
    qn = np.linalg.norm(Q, axis=0)
    assert np.allclose(mcn, np.where(qn > 1, Q / np.where(qn > 1, qn, 1.0), Q))


def test_maxcolnorm_threshold_is_one():
    # src/cpd/constraints/maxcolnorm.c:32-61: a column is divided by its norm
    # only when that norm exceeds 1; columns of norm in (0, 1] are untouched
    # whatever their size (0.3, 0.55, 0.7, 0.95), those above land on norm 1
    rng = np.random.default_rng(14)
    P = rng.standard_normal((40, 6))
    want_norms = np.array([0.3, 0.55, 0.7, 0.95, 1.2, 4.0])
    P = P / np.linalg.norm(P, axis=0) * want_norms
    for kind in (1, 2):
        Q = np.where(P < 0, 0.0, P) if kind == 2 else P
        qn = np.linalg.norm(Q, axis=0)
        got = orc.prox(kind, P)
        below = qn <= 1.0
        assert np.array_equal(got[:, below], Q[:, below]), kind
        assert np.allclose(np.linalg.norm(got[:, ~below], axis=0), 1.0, rtol=1e-14)
        assert np.allclose(got[:, ~below], Q[:, ~below] / qn[~below], rtol=1e-15, atol=0)
    # a max-column-norm ADMM whose least-squares solution has column norms
    # between 0.5 and 1 converges to that solution (the constraint is inactive)
    R, I = 4, 30
    G = _spd(R, 15, scale=2.0)
    ls = rng.standard_normal((I, R))
    ls = ls / np.linalg.norm(ls, axis=0) * np.array([0.6, 0.75, 0.9, 0.98])
    M = ls @ G
    assert np.all((np.linalg.norm(ls, axis=0) > 0.5) & (np.linalg.norm(ls, axis=0) < 1))
    P1, _, _, _ = orc.admm(1, M, np.zeros((I, R)), np.zeros((I, R)), G, 3000, 0.0)
    assert np.allclose(P1, ls, rtol=1e-8, atol=1e-11)


def test_admm_identity_prox_converges_to_solve():
    R, I = 5, 40
    G = _spd(R, 1)
    rng = np.random.default_rng(5)
    M = rng.standard_normal((I, R))
    P0 = rng.standard_normal((I, R))
    P, D, its, rho = orc.admm(0, M, P0, np.zeros((I, R)), G, 400, 0.0)
    # with tolerance 0 it stops only at an exact floating-point fixed point
    assert 1 < its <= 400
    assert abs(rho - np.trace(G) / R) <= 1e-14 * rho
    assert np.allclose(P, np.linalg.solve(G, M.T).T, rtol=1e-8, atol=1e-10)
    assert np.max(np.abs(D)) <= 1e-12


def test_admm_nonneg_kkt():
    R, I = 6, 30
    G = _spd(R, 2)
    rng = np.random.default_rng(6)
    M = rng.standard_normal((I, R)) * 3.0
    P, D, its, rho = orc.admm(3, M, np.zeros((I, R)), np.zeros((I, R)), G, 3000, 0.0)
    assert np.all(P >= 0)
    g = P @ G - M           # gradient of 1/2 p^T G p - m^T p per row
    scale = np.max(np.abs(M))
    assert np.all(g[P == 0] >= -1e-6 * scale)
    assert np.max(np.abs(g[P > 1e-9])) <= 1e-6 * scale


def test_admm_maxcolnorm_inactive_and_active():
    R, I = 4, 25
    G = _spd(R, 3, scale=5.0)
    rng = np.random.default_rng(7)
    M = rng.standard_normal((I, R)) * 0.01
    ls = np.linalg.solve(G, M.T).T
    assert np.all(np.linalg.norm(ls, axis=0) < 1)
    P, D, its, rho = orc.admm(1, M, np.zeros((I, R)), np.zeros((I, R)), G, 2000, 0.0)
    assert np.allclose(P, ls, rtol=1e-7, atol=1e-10)
    M2 = M * 1e4
    P2, D2, its2, rho2 = orc.admm(1, M2, np.zeros((I, R)), np.zeros((I, R)), G, 2000, 0.0)
    assert np.all(np.linalg.norm(P2, axis=0) <= 1 + 1e-12)


def test_admm_stops_on_tolerance():
    R, I = 4, 20
    G = _spd(R, 8)
    rng = np.random.default_rng(9)
    M = rng.standard_normal((I, R))
    P, D, its, rho = orc.admm(3, M, np.zeros((I, R)), np.zeros((I, R)), G, 50, 1e-2)
    assert 1 <= its < 50


@pytest.mark.parametrize("chunk", [7, 10, 50, 1000])
def test_admm_chunked_is_independent_chunks(chunk):
    """Each chunk of rows (the last ragged) runs its own ADMM to its own
    convergence from the same G + rho I; a chunk at least I rows long is the
    one-block ADMM; the count is iterations x rows summed over chunks."""
    R, I = 5, 43
    G = _spd(R, 12)
    rng = np.random.default_rng(13)
    M = rng.standard_normal((I, R)) * 2.0
    P0 = np.abs(rng.standard_normal((I, R)))
    D0 = rng.standard_normal((I, R)) * 0.1
    P, D, tot = orc.admm_chunked(3, M, P0, D0, G, chunk, 50, 1e-2)
    want_tot = 0
    for s0 in range(0, I, chunk):
        s1 = min(s0 + chunk, I)
        Pc, Dc, its, rho = orc.admm(3, M[s0:s1], P0[s0:s1], D0[s0:s1], G, 50, 1e-2)
        assert np.array_equal(P[s0:s1], Pc) and np.array_equal(D[s0:s1], Dc)
        want_tot += its * (s1 - s0)
    assert tot == want_tot
    assert np.all(P >= 0)
    if chunk >= I:
        Pf, Df, its, rho = orc.admm(3, M, P0, D0, G, 50, 1e-2)
        assert np.array_equal(P, Pf) and tot == its * I


def test_cpd_nonneg_chunked_kkt():
    """Nonneg CPD with chunks of 4 rows: a few sweeps keep every factor
    non-negative and each chunk's ADMM (given enough inner iterations and a
    tight tolerance) reaches the nonneg least-squares optimum of its rows."""
    dims = [9, 8, 7]
    inds, vals = _tensor(dims, 200, 21)
    c = orc.Coo.from_arrays(inds, vals, dims)
    csfs = [orc.Csf(c, orc.mode_order(dims, 3, m)) for m in range(3)]
    init = [np.abs(f) for f in _init(dims, 3, 23)]
    cons = [(2, 3, 0.0, 4)] * 3
    f, lam, fits = orc.cpd_als_con(csfs, [0, 1, 2], init, 3, 0.0, cons, False,
                                   max_inner=5000, inner_tol=0.0)
    assert all(np.all(x >= 0) for x in f)
    # the last mode's update: nonneg least squares against the other factors
    G = (f[0].T @ f[0]) * (f[1].T @ f[1])
    X = _dense(inds, vals, dims)
    Mm = np.einsum("ijk,ir,jr->kr", X, f[0], f[1])
    g = f[2] @ G - Mm
    scale = np.max(np.abs(Mm))
    assert np.all(g[f[2] == 0] >= -1e-6 * scale)
    assert np.max(np.abs(g[f[2] > 1e-9])) <= 1e-6 * scale


def test_cpd_frob_closed_form():
    dims = [8, 6, 7]
    inds, vals = _tensor(dims, 150, 3)
    c = orc.Coo.from_arrays(inds, vals, dims)
    csfs = [orc.Csf(c, orc.mode_order(dims, 3, m)) for m in range(3)]
    init = _init(dims, 3, 17)
    cons0 = [(1, 0, 0.0)] * 3
    f0, l0, fits0 = orc.cpd_als_con(csfs, [0, 1, 2], init, 1, 0.0, cons0, False)
    # mu = 0: a plain normal-equation solve of every mode, lambda stays 1
    assert np.allclose(l0, 1.0)
    start = [f / np.linalg.norm(f, axis=0) for f in init]
    X = _dense(inds, vals, dims)
    cur = [s.copy() for s in start]
    for m in range(3):
        others = [cur[k] for k in range(3) if k != m]
        G = np.ones((3, 3))
        for o in others:
            G *= o.T @ o
        sub = "ijk," + ",".join("%sr" % "ijk"[k] for k in range(3) if k != m) + \
              "->%sr" % "ijk"[m]
        Mm = np.einsum(sub, X, *others)
        cur[m] = np.linalg.solve(G, Mm.T).T
    for m in range(3):
        assert np.allclose(f0[m], cur[m], rtol=1e-9, atol=1e-9)
    # mu > 0 on mode 1: (G + mu I) U1^T = M1^T for the final sweep's inputs
    mu = 0.5
    cons = [(1, 0, 0.0), (1, 0, mu), (1, 0, 0.0)]
    f1, l1, fits1 = orc.cpd_als_con(csfs, [0, 1, 2], init, 1, 0.0, cons, False)
    # mode 1 is solved after mode 0, before mode 2: its Gram uses the new U0
    # and the starting U2
    G = (f1[0].T @ f1[0]) * (start[2].T @ start[2]) + mu * np.eye(3)
    M1 = np.einsum("ijk,ir,kr->jr", X, f1[0], start[2])
    assert np.allclose(f1[1] @ G, M1, rtol=1e-9, atol=1e-9)


def test_lasso_prox_is_the_l1_proximal_point():
    # x = prox(v) minimises (x - v)^2 / 2 + thr |x|: either x != 0 and
    # x - v + thr sign(x) = 0, or x = 0 and |v| <= thr
    rng = np.random.default_rng(5)
    v = rng.standard_normal((40, 7)) * 3.0
    v[0, :3] = [1.5, -1.5, 0.0]                      # exactly at the threshold
    thr = 1.5
    x = orc.prox(4, v, thr)
    nz = x != 0
    assert np.all(np.abs(x[nz] - v[nz] + thr * np.sign(x[nz])) <= 1e-15 * np.abs(v[nz]).max())
    assert np.all(np.abs(v[~nz]) <= thr)
    assert np.all(np.sign(x[nz]) == np.sign(v[nz]))
    assert 0 < nz.sum() < nz.size
    # threshold 0 is the identity
    assert np.array_equal(orc.prox(4, v, 0.0), v)


def test_admm_lasso_meets_l1_optimality():
    # ADMM (src/cpd/admm.c:240-308) with the lasso prox solves
    # min_P tr(P G P^T) / 2 - tr(M^T P) + lam |P|_1: at convergence
    # E = M - P G equals lam sign(P) where P != 0 and |E| <= lam where P = 0;
    # lam = 0 reproduces the unconstrained run exactly
    rng = np.random.default_rng(11)
    I, R, lam = 60, 5, 2.0
    A = rng.standard_normal((15, R))
    G = A.T @ A + 0.5 * np.eye(R)
    M = rng.standard_normal((I, R)) * 4.0
    Z = np.zeros((I, R))
    P, D, its, rho = orc.admm(4, M, Z, Z, G, 3000, 1e-14, lam=lam)
    E = M - P @ G
    nz = P != 0
    assert 0 < nz.sum() < nz.size
    assert np.max(np.abs(E[nz] - lam * np.sign(P[nz]))) <= 1e-5 * lam
    assert np.max(np.abs(E[~nz])) <= lam * (1 + 1e-5)
    P0, D0, its0, _ = orc.admm(0, M, Z, Z, G, 40, 1e-3)
    P4, D4, its4, _ = orc.admm(4, M, Z, Z, G, 40, 1e-3, lam=0.0)
    assert its0 == its4 and np.array_equal(P0, P4) and np.array_equal(D0, D4)


def test_ntf_proxes_are_nonneg_proximal_points():
    # nonneg_frob.c / nonneg_lasso.c: with step 1/rho, x = prox(v) minimises
    # rho/2 (x - v)^2 + lam/2 x^2 (frob) or rho/2 (x - v)^2 + lam x (lasso)
    # over x >= 0; the KKT conditions of that scalar problem are checked on
    # every entry, and a brute-force grid minimum agrees
    rng = np.random.default_rng(17)
    v = rng.standard_normal((30, 6)) * 2.0
    lam, rho = 0.7, 1.9
    xf = orc.prox(8, v, rho / (lam + rho))
    xl = orc.prox(9, v, lam / rho)
    for x, g in ((xf, rho * (xf - v) + lam * xf), (xl, rho * (xl - v) + lam)):
        assert np.all(x >= 0)
        pos = x > 0
        assert np.max(np.abs(g[pos])) <= 1e-13 * np.abs(v).max() * rho
        assert np.all(g[~pos] >= -1e-13)
        assert 0 < pos.sum() < pos.size
    grid = np.linspace(0.0, 8.0, 800001)
    for (i, j) in ((0, 0), (3, 2), (7, 5)):
        f = rho / 2 * (grid - v[i, j]) ** 2 + lam / 2 * grid ** 2
        assert abs(grid[np.argmin(f)] - xf[i, j]) <= 2e-5
        f = rho / 2 * (grid - v[i, j]) ** 2 + lam * grid
        assert abs(grid[np.argmin(f)] - xl[i, j]) <= 2e-5


@pytest.mark.parametrize("prox", [8, 9])
def test_admm_ntf_meets_kkt(prox):
    # ADMM with these proxes solves min_{P >= 0} tr(P G P^T)/2 - tr(M^T P)
    # + lam/2 |P|^2 (8) or + lam sum(P) (9): per entry the gradient
    # g = P G - M + lam P (8) or + lam (9) is 0 where P > 0, >= 0 where P = 0
    rng = np.random.default_rng(19 + prox)
    I, R, lam = 40, 5, 1.5
    A = rng.standard_normal((15, R))
    G = A.T @ A + 0.5 * np.eye(R)
    M = rng.standard_normal((I, R)) * 4.0
    Z = np.zeros((I, R))
    P, D, its, rho = orc.admm(prox, M, Z, Z, G, 5000, 1e-15, lam=lam)
    g = P @ G - M + (lam * P if prox == 8 else lam)
    pos = P > 0
    assert np.all(P >= 0) and 0 < pos.sum() < pos.size
    assert np.max(np.abs(g[pos])) <= 1e-5 * np.abs(M).max()
    assert np.all(g[~pos] >= -1e-5 * np.abs(M).max())


def test_rowsimp_prox_is_the_simplex_projection():
    # the Euclidean projection of v onto {x >= 0, sum x = 1} is
    # max(v - theta, 0) with theta the root of sum max(v - theta, 0) = 1;
    # theta is found here by bisection, independently of the sort
    rng = np.random.default_rng(8)
    V = rng.standard_normal((30, 6)) * 2.0
    V[0] = [0.3, 0.3, 0.3, 0.3, -1.0, 2.0]             # ties
    V[1] = [5.0, -3.0, 0.1, 0.2, 0.0, 0.0]             # one active entry
    X = orc.prox(5, V)
    for v, x in zip(V, X):
        lo, hi = v.min() - 1.0, v.max()
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if np.maximum(v - mid, 0.0).sum() > 1.0:
                lo = mid
            else:
                hi = mid
        want = np.maximum(v - 0.5 * (lo + hi), 0.0)
        assert np.all(x >= 0) and abs(x.sum() - 1.0) <= 1e-12
        assert np.allclose(x, want, atol=1e-12)
    assert np.array_equal(orc.prox(5, np.array([[4.0]])), np.array([[1.0]]))


def _smooth_matrix(I):
    # src/cpd/constraints/smooth.c:30-79 (p_form_banded) without lam and rho:
    # stencil [1 -4 6 -4 1], 5 on the first and last diagonal entries
    K = np.zeros((I, I))
    for i in range(I):
        K[i, i] = 5.0 if i in (0, I - 1) else 6.0
        if i + 1 < I:
            K[i, i + 1] = K[i + 1, i] = -4.0
        if i + 2 < I:
            K[i, i + 2] = K[i + 2, i] = 1.0
    return K


@pytest.mark.parametrize("I", [1, 2, 3, 4, 7, 40])
def test_smooth_prox_is_the_banded_solve(I):
    # every column x becomes rho (lam K + rho I)^{-1} x
    rng = np.random.default_rng(I)
    P = rng.standard_normal((I, 3))
    lam, rho = 0.7, 1.3
    want = rho * np.linalg.solve(lam * _smooth_matrix(I) + rho * np.eye(I), P)
    assert np.allclose(orc.smooth_prox(P, lam, rho), want, rtol=1e-12, atol=1e-13)


def test_admm_smooth_meets_optimality():
    # ADMM with the smoothness prox solves
    # min_P tr(P G P^T) / 2 - tr(M^T P) + lam / 2 sum_c P_c^T K P_c, whose
    # minimiser satisfies P G + lam K P = M: a linear system solved densely
    rng = np.random.default_rng(13)
    I, R, lam = 25, 4, 0.8
    A = rng.standard_normal((12, R))
    G = A.T @ A + 0.5 * np.eye(R)
    M = rng.standard_normal((I, R)) * 3.0
    Z = np.zeros((I, R))
    P, D, its, rho = orc.admm(6, M, Z, Z, G, 5000, 1e-20, lam=lam)
    K = _smooth_matrix(I)
    big = np.kron(G, np.eye(I)) + np.kron(np.eye(R), lam * K)
    want = np.linalg.solve(big, M.T.reshape(-1)).reshape(R, I).T
    assert np.max(np.abs(P - want)) <= 1e-6 * np.max(np.abs(want))


def test_colsimp_prox_projects_every_column():
    # columns onto the simplex: the row projection of the transpose, and each
    # column sums to 1 with max(v - theta, 0) for the bisection theta
    rng = np.random.default_rng(21)
    V = rng.standard_normal((40, 5)) * 0.3
    X = orc.prox(7, V)
    assert np.array_equal(X, orc.prox(5, V.T).T)
    for v, x in zip(V.T, X.T):
        lo, hi = v.min() - 1.0, v.max()
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if np.maximum(v - mid, 0.0).sum() > 1.0:
                lo = mid
            else:
                hi = mid
        assert abs(x.sum() - 1.0) <= 1e-12
        assert np.allclose(x, np.maximum(v - 0.5 * (lo + hi), 0.0), atol=1e-12)

