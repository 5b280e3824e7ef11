"""GPU parity of the dense CPD steps and of splatt_cpd against the oracle.

Dense kernels (device arrays through the exported spd_dense_* entries):
  Gram         X^T X on MFMA against numpy, with and without the in-place
               column division and the FP32 copy; ranks on both sides of 64.
  chol_inv     ginv G = I for the Hadamard of Grams (+ Frobenius term,
               + rho = trace/R for ADMM); a non-positive pivot is reported.
  apply        M ginv against numpy's solve, FP64 and FP32 right-hand sides,
               and the fused column norms.
  admm_step    one inner iteration against the oracle's or_admm with
               max_inner = 1, for every proximal operator.

splatt_cpd (C API; factors from splatt_alloc_cpd after srand):
  FP64: fit per run, lambda and factors against or_cpd_als_con from the same
  start (relative 1e-8 on the fit; the two differ only in rounding:
  explicit inverse vs triangular solves, summation order).
  FP32 MTTKRP: fit within 1e-4 of the FP64 oracle.
  Frobenius on one mode, and ADMM with max-column-norm constraints.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as orc
import pysplatt as ps
from conftest import tensor_path

pytestmark = pytest.mark.gpu


def _dev(a):
    return ps.DevBuf.from_host(np.ascontiguousarray(a))


def _rp(R):
    return (R + 15) // 16 * 16


@pytest.mark.parametrize("R", [1, 5, 16, 32, 40, 64, 100])
@pytest.mark.parametrize("I", [1, 37, 1000])
def test_gram(gpu_available, I, R):
    L = ps.lib()
    rng = np.random.default_rng(I * 131 + R)
    X = rng.standard_normal((I, R))
    dX = _dev(X)
    out = ps.DevBuf(R * R * 8)
    assert L.spd_dense_gram(dX.ptr, I, R, None, None, out.ptr) == 0
    G = out.to_host((R, R), np.float64)
    want = X.T @ X
    assert np.allclose(G, want, rtol=1e-12, atol=1e-12 * np.max(np.abs(want)))
    assert np.array_equal(G, G.T)
    # with column division and FP32 copy
    div = np.abs(rng.standard_normal(R)) + 0.5
    dd = _dev(div)
    x32 = ps.DevBuf(I * R * 4)
    assert L.spd_dense_gram(dX.ptr, I, R, dd.ptr, x32.ptr, out.ptr) == 0
    Xd = dX.to_host((I, R), np.float64)
    assert np.array_equal(Xd, X / div)
    assert np.array_equal(x32.to_host((I, R), np.float32), (X / div).astype(np.float32))
    G2 = out.to_host((R, R), np.float64)
    want2 = (X / div).T @ (X / div)
    assert np.allclose(G2, want2, rtol=1e-12, atol=1e-12 * np.max(np.abs(want2)))


def _grams(R, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        A = rng.standard_normal((2 * R + 3, R))
        out.append(A.T @ A)
    return out


@pytest.mark.parametrize("R", [3, 16, 32, 33, 48, 64, 80])
@pytest.mark.parametrize("admm", [0, 1])
def test_chol_inv(gpu_available, R, admm):
    L = ps.lib()
    N = 3
    grams = _grams(R, N, R + 7 * admm)
    bufs = [_dev(g) for g in grams]
    arr = (C.c_void_p * N)(*[b.ptr for b in bufs])
    RP = _rp(R)
    ginv = ps.DevBuf(RP * RP * 8)
    rho = ps.DevBuf(8)
    info = ps.DevBuf(8)
    mode, mu = 1, 0.25
    assert L.spd_dense_chol_inv(arr, N, mode, R, mu, admm, ginv.ptr, rho.ptr, info.ptr) == 0
    G = np.ones((R, R))
    for m in range(N):
        if m != mode:
            G *= np.triu(grams[m]) + np.triu(grams[m], 1).T
    G += mu * np.eye(R)
    if admm:
        r = rho.to_host((1,), np.float64)[0]
        assert abs(r - np.trace(G) / R) <= 1e-13 * abs(r)
        G += r * np.eye(R)
    Gi = ginv.to_host((RP, RP), np.float64)
    assert np.all(Gi[R:, :] == 0) and np.all(Gi[:, R:] == 0)
    Gi = Gi[:R, :R]
    assert np.array_equal(Gi, Gi.T)
    err = np.max(np.abs(Gi @ G - np.eye(R)))
    assert err <= 1e-9, err
    assert info.to_host((1,), np.int32)[0] == 0


def test_chol_reports_bad_pivot(gpu_available):
    L = ps.lib()
    R = 4
    g = -np.eye(R)
    bufs = [_dev(np.eye(R)), _dev(g)]
    arr = (C.c_void_p * 2)(*[b.ptr for b in bufs])
    ginv = ps.DevBuf(16 * 16 * 8)
    rho = ps.DevBuf(8)
    info = ps.DevBuf(8)
    assert L.spd_dense_chol_inv(arr, 2, 0, R, 0.0, 0, ginv.ptr, rho.ptr, info.ptr) == 0
    assert info.to_host((1,), np.int32)[0] == 1


def _ld_cholesky_solve(G, M):
    """X = M G^{-1} in long double (64-bit mantissa) by a textbook Cholesky:
    the near-exact solution that the FP64 paths are measured against."""
    G = np.asarray(G, dtype=np.longdouble)
    R = G.shape[0]
    Lc = np.zeros((R, R), dtype=np.longdouble)
    for j in range(R):
        d = G[j, j] - np.dot(Lc[j, :j], Lc[j, :j])
        Lc[j, j] = np.sqrt(d)
        for i in range(j + 1, R):
            Lc[i, j] = (G[i, j] - np.dot(Lc[i, :j], Lc[j, :j])) / Lc[j, j]
    B = np.asarray(M, dtype=np.longdouble).T.copy()
    for i in range(R):
        B[i] = (B[i] - Lc[i, :i] @ B[:i]) / Lc[i, i]
    for i in range(R - 1, -1, -1):
        B[i] = (B[i] - Lc[i + 1:, i] @ B[i + 1:]) / Lc[i, i]
    return B.T


@pytest.mark.parametrize("R", [16, 32, 64])
@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6, 1e8, 1e10])
def test_gj_inverse_matches_cholesky_when_ill_conditioned(gpu_available, R, kappa):
    # The normal equations are solved as M G^{-1} with an explicit
    # Gauss-Jordan inverse instead of the reference's potrf/potrs
    # (src/matrix.c:104-151). On SPD Grams with condition numbers 1e2..1e10
    # its forward error is held to that of LAPACK's Cholesky solve (scipy's
    # dpotrf/dpotrs, the reference's routines), both measured against a
    # long-double solution: within 20x of it (plus 1e-14), and within the
    # backward-stable bound R * kappa * eps * 50. After a non-positive pivot
    # both the reference (which prints dpotrf's info and continues) and this
    # library (info = the 1-based column, as dpotrf) leave the solution
    # undefined; test_chol_reports_bad_pivot checks the status.
    import scipy.linalg as sla
    L = ps.lib()
    rng = np.random.default_rng(int(R * 1000 + np.log10(kappa)))
    Q, _ = np.linalg.qr(rng.standard_normal((R, R)))
    lam = np.geomspace(1.0, 1.0 / kappa, R)
    G = (Q * lam) @ Q.T
    G = 0.5 * (G + G.T)
    I = 2000
    M = rng.standard_normal((I, R))
    bufs = [_dev(G), _dev(np.ones((R, R)))]
    arr = (C.c_void_p * 2)(*[b.ptr for b in bufs])
    RP = _rp(R)
    ginv = ps.DevBuf(RP * RP * 8)
    rho = ps.DevBuf(8)
    info = ps.DevBuf(8)
    assert L.spd_dense_chol_inv(arr, 2, 1, R, 0.0, 0, ginv.ptr, rho.ptr, info.ptr) == 0
    assert info.to_host((1,), np.int32)[0] == 0
    dM = _dev(M)
    X = ps.DevBuf(I * R * 8)
    cn = ps.DevBuf(R * 8)
    assert L.spd_dense_apply(dM.ptr, 0, ginv.ptr, I, R, X.ptr, cn.ptr) == 0
    got = X.to_host((I, R), np.float64)
    chol = sla.cho_solve(sla.cho_factor(G, lower=True), M.T).T
    exact = _ld_cholesky_solve(G, M)
    nrm = float(np.linalg.norm(exact.astype(np.float64)))
    err_gj = float(np.linalg.norm((got - exact).astype(np.float64))) / nrm
    err_ch = float(np.linalg.norm((chol - exact).astype(np.float64))) / nrm
    print("R %d kappa %.0e: GJ %.2e Cholesky %.2e" % (R, kappa, err_gj, err_ch))
    assert err_gj <= 20.0 * err_ch + 1e-14, (err_gj, err_ch)
    assert err_gj <= 50.0 * R * kappa * 2.2e-16, err_gj


@pytest.mark.parametrize("R", [2, 16, 24, 48, 64, 70])
@pytest.mark.parametrize("f32", [False, True])
def test_apply(gpu_available, R, f32):
    L = ps.lib()
    I = 777
    rng = np.random.default_rng(R)
    G = _grams(R, 1, R)[0] + np.eye(R)
    bufs = [_dev(G), _dev(np.ones((R, R)))]
    arr = (C.c_void_p * 2)(*[b.ptr for b in bufs])
    RP = _rp(R)
    ginv = ps.DevBuf(RP * RP * 8)
    rho = ps.DevBuf(8)
    info = ps.DevBuf(8)
    assert L.spd_dense_chol_inv(arr, 2, 1, R, 0.0, 0, ginv.ptr, rho.ptr, info.ptr) == 0
    M = rng.standard_normal((I, R))
    if f32:
        M = M.astype(np.float32)
    dM = _dev(M)
    X = ps.DevBuf(I * R * 8)
    cn = ps.DevBuf(R * 8)
    assert L.spd_dense_apply(dM.ptr, 1 if f32 else 0, ginv.ptr, I, R, X.ptr, cn.ptr) == 0
    got = X.to_host((I, R), np.float64)
    want = np.linalg.solve(G, M.astype(np.float64).T).T
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9 * np.max(np.abs(want)))
    norms = cn.to_host((R,), np.float64)
    assert np.allclose(norms, np.linalg.norm(got, axis=0), rtol=1e-12)


@pytest.mark.parametrize("R", [8, 32, 70])
@pytest.mark.parametrize("prox", [0, 1, 2, 3])
def test_admm_step_matches_oracle(gpu_available, R, prox):
    L = ps.lib()
    I = 300
    rng = np.random.default_rng(R * 10 + prox)
    grams = _grams(R, 3, R + prox)
    mode = 2
    Gfull = np.ones((R, R))
    for m in range(3):
        if m != mode:
            Gfull *= grams[m]
    M = rng.standard_normal((I, R)) * 5
    P0 = rng.standard_normal((I, R))
    D0 = rng.standard_normal((I, R)) * 0.1
    wantP, wantD, its, wrho = orc.admm(prox, M, P0, D0, Gfull, 1, 0.0)
    bufs = [_dev(g) for g in grams]
    arr = (C.c_void_p * 3)(*[b.ptr for b in bufs])
    RP = _rp(R)
    ginv = ps.DevBuf(RP * RP * 8)
    rho = ps.DevBuf(8)
    info = ps.DevBuf(8)
    assert L.spd_dense_chol_inv(arr, 3, mode, R, 0.0, 1, ginv.ptr, rho.ptr, info.ptr) == 0
    dM, dP, dD = _dev(M), _dev(P0), _dev(D0)
    aux = ps.DevBuf(I * R * 8)
    init = ps.DevBuf(I * R * 8)
    sums = ps.DevBuf((4 + R) * 8)
    assert L.spd_dense_admm_step(dM.ptr, 0, ginv.ptr, rho.ptr, I, R, dP.ptr, aux.ptr,
                                 dD.ptr, init.ptr, prox, sums.ptr) == 0
    P = dP.to_host((I, R), np.float64)
    D = dD.to_host((I, R), np.float64)
    s = sums.to_host((4 + R,), np.float64)
    assert abs(rho.to_host((1,), np.float64)[0] - wrho) <= 1e-13 * wrho
    tol = 1e-8
    assert np.allclose(P, wantP, rtol=tol, atol=tol * np.max(np.abs(wantP)))
    # D is a difference of O(|P|) terms (zero up to rounding for prox 0)
    assert np.allclose(D, wantD, rtol=tol, atol=tol * np.max(np.abs(wantP)))
    assert np.array_equal(init.to_host((I, R), np.float64), P0)
    A = aux.to_host((I, R), np.float64)
    assert np.allclose(s[:4], [np.sum(P * P), np.sum((P - A) ** 2), np.sum((P - P0) ** 2),
                               np.sum(D * D)], rtol=1e-12)


@pytest.mark.parametrize("R,I", [(5, 1), (16, 37), (24, 700), (32, 2000), (33, 500),
                                 (64, 3000)])
@pytest.mark.parametrize("prox", [0, 1, 2, 3, 4, 6, 7, 8, 9])
@pytest.mark.parametrize("tol,max_inner", [(1e-2, 50), (7e-2, 25), (0.0, 6)])
def test_admm_solve_matches_oracle(gpu_available, R, I, prox, tol, max_inner):
    # the whole inner loop: same iterations, P and D at 1e-8 (prox 4: lasso
    # with multiplier 2; prox 6: column smoothness with multiplier 0.5;
    # prox 7: column simplex; 8: non-negative frob 0.5; 9: non-negative L1
    # 2). Ranks 17-32 (24, 32: the streaming workload's rank) run the
    # two-tile pass.
    rng = np.random.default_rng(R * 7 + prox + int(tol * 1000))
    A = rng.standard_normal((3 * R, R))
    G = A.T @ A + 0.5 * np.eye(R)
    M = rng.standard_normal((I, R)) * 4.0
    P0 = rng.standard_normal((I, R))
    D0 = rng.standard_normal((I, R)) * 0.1
    lam = {4: 2.0, 6: 0.5, 8: 0.5, 9: 2.0}.get(prox, 0.0)
    wantP, wantD, wits, wrho = orc.admm(prox, M, P0, D0, G, max_inner, tol, lam=lam)
    L = ps.lib()
    RP = _rp(R)
    bufs = [_dev(G), _dev(np.ones((R, R)))]
    arr = (C.c_void_p * 2)(*[b.ptr for b in bufs])
    ginv = ps.DevBuf(RP * RP * 8)
    rho = ps.DevBuf(8)
    info = ps.DevBuf(8)
    assert L.spd_dense_chol_inv(arr, 2, 1, R, 0.0, 1, ginv.ptr, rho.ptr, info.ptr) == 0
    dM, dP, dD = _dev(M), _dev(P0), _dev(D0)
    aux = ps.DevBuf(I * R * 8)
    init = ps.DevBuf(I * R * 8)
    sums = ps.DevBuf((4 + R) * 8)
    ctl = ps.DevBuf(8)
    assert L.spd_dense_admm_solve(dM.ptr, 0, ginv.ptr, rho.ptr, I, R, dP.ptr, dD.ptr, aux.ptr,
                                  init.ptr, prox, C.c_double(lam), max_inner, C.c_double(tol),
                                  sums.ptr, ctl.ptr) == 0
    c = ctl.to_host((2,), np.int32)
    assert int(c[1]) == wits
    P = dP.to_host((I, R), np.float64)
    D = dD.to_host((I, R), np.float64)
    t = 1e-8
    scale = max(np.max(np.abs(wantP)), np.max(np.abs(wantD)))
    assert np.allclose(P, wantP, rtol=t, atol=t * scale)
    assert np.allclose(D, wantD, rtol=t, atol=t * scale)


# ---- splatt_cpd ---------------------------------------------------------------

def _twomode(path=None, coo=None, alloc=ps.CSF_TWOMODE):
    opts = ps.Opts(csf_alloc=alloc)
    if path is not None:
        nm = C.c_uint64()
        out = C.POINTER(ps.SplattCsf)()
        assert ps.lib().splatt_csf_load(path.encode(), C.byref(nm), C.byref(out),
                                        opts.ptr) == ps.SUCCESS
        return ps.CsfSet(out, opts)
    inds, vals = coo
    return ps.CsfSet.convert(inds, vals, opts)


def _oracle_from_path(path):
    coo = orc.Coo.read(path)
    coo.remove_empty()
    return coo


def _run_both(path, R, its, cons=None, unconstrained=True, configure=None,
              max_inner=50, inner_tol=1e-2, f32=False, seed=7):
    cs = _twomode(path=path)
    L = ps.lib()
    L.splatt_amd_set_mttkrp_precision(1 if f32 else 0)
    try:
        res = ps.run_cpd(cs, R, max_its=its, tol=0.0, seed=seed, configure=configure,
                         max_inner=max_inner, inner_tol=inner_tol)
    finally:
        L.splatt_amd_set_mttkrp_precision(0)
    assert res["rc"] == ps.SUCCESS, ps.last_error()
    coo = _oracle_from_path(path)
    dims = res["dims"]
    N = len(dims)
    a0 = cs.arrays(0)
    perm0 = a0["perm"]
    ocsfs = [orc.Csf(coo, perm0)]
    if cs.count >= 2:
        ocsfs.append(orc.Csf(coo, cs.arrays(1)["perm"]))
    mode_csf = [1 if (cs.count >= 2 and perm0[-1] == m) else 0 for m in range(N)]
    if cons is None:
        cons = [(0, 0, 0.0)] * N
    facs, lam, fits = orc.cpd_als_con(ocsfs, mode_csf, res["init"], its, 0.0, cons,
                                      unconstrained, max_inner, inner_tol)
    return res, facs, lam, fits


def _model_close(res, facs, lam, rtol):
    """Compare the two Kruskal models column by column (scaled factors)."""
    for m in range(len(facs)):
        a = res["factors"][m] * (res["lam"] if m == 0 else 1.0)
        b = facs[m] * (lam if m == 0 else 1.0)
        scale = np.max(np.abs(b))
        assert np.allclose(a, b, rtol=rtol, atol=rtol * scale), m


@pytest.mark.parametrize("name,R", [("small.tns", 4), ("med.tns", 8), ("med4.tns", 6),
                                    ("med5.tns", 5)])
def test_cpd_fp64_matches_oracle(gpu_available, name, R):
    res, facs, lam, fits = _run_both(tensor_path(name), R, 5)
    assert abs(res["fit"] - fits[-1]) <= 1e-8 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-6)


def test_cpd_fp32_fit_within_1e4(gpu_available):
    res, facs, lam, fits = _run_both(tensor_path("med.tns"), 8, 10, f32=True)
    assert abs(res["fit"] - fits[-1]) <= 1e-4


def test_cpd_frob_one_mode(gpu_available):
    def configure(L, co):
        modes = np.array([1], dtype=np.uint64)
        L.splatt_register_frob(co, 0.1, modes.ctypes.data_as(ps.u64p), 1)
    res, facs, lam, fits = _run_both(tensor_path("med.tns"), 6, 4,
                                     cons=[(1, 0, 0.0), (1, 0, 0.1), (1, 0, 0.0)],
                                     unconstrained=False, configure=configure)
    assert np.allclose(res["lam"], 1.0)
    assert abs(res["fit"] - fits[-1]) <= 1e-8 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-6)


def test_cpd_admm_maxcolnorm(gpu_available):
    def configure(L, co):
        modes = np.array([0, 2], dtype=np.uint64)
        L.splatt_register_maxcolnorm(co, modes.ctypes.data_as(ps.u64p), 2)
    res, facs, lam, fits = _run_both(tensor_path("med.tns"), 5, 3,
                                     cons=[(2, 1, 0.0), (1, 0, 0.0), (2, 1, 0.0)],
                                     unconstrained=False, configure=configure,
                                     max_inner=25, inner_tol=7e-2)
    for m in (0, 2):
        assert np.all(np.linalg.norm(res["factors"][m], axis=0) <= 1 + 1e-9)
    assert abs(res["fit"] - fits[-1]) <= 1e-6 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-5)


@pytest.mark.parametrize("chunk", [50, 7, 0])
def test_cpd_nonneg_chunked(gpu_available, chunk):
    # nonneg on modes 0 and 2 with the reference's default of 50 rows per
    # ADMM chunk, a ragged 7, and 0 (one block); each chunk converges on its
    # own, so the iteration counts of the chunks differ
    def configure(L, co):
        modes = np.array([0, 2], dtype=np.uint64)
        L.splatt_register_nonneg(co, modes.ctypes.data_as(ps.u64p), 2)
        for m in range(3):
            co[0].chunk_sizes[m] = chunk
    res, facs, lam, fits = _run_both(tensor_path("med.tns"), 5, 3,
                                     cons=[(2, 3, 0.0, chunk), (1, 0, 0.0), (2, 3, 0.0, chunk)],
                                     unconstrained=False, configure=configure)
    for m in (0, 2):
        assert np.all(res["factors"][m] >= 0)
    assert abs(res["fit"] - fits[-1]) <= 1e-6 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-5)


@pytest.mark.parametrize("chunk", [50, 0])
def test_cpd_lasso_matches_oracle(gpu_available, chunk):
    # splatt_register_lasso on modes 0 and 2 (multiplier 1e-4 zeroes part of
    # both factors), in chunks of 50 rows (the reference's default) or one
    # block
    lam_l1 = 1e-4

    def configure(L, co):
        modes = np.array([0, 2], dtype=np.uint64)
        assert L.splatt_register_lasso(co, lam_l1, modes.ctypes.data_as(ps.u64p), 2) == \
            ps.SUCCESS
        for m in range(3):
            co[0].chunk_sizes[m] = chunk
    res, facs, lam, fits = _run_both(tensor_path("med.tns"), 5, 3,
                                     cons=[(2, 4, 0.0, chunk, lam_l1), (1, 0, 0.0),
                                           (2, 4, 0.0, chunk, lam_l1)],
                                     unconstrained=False, configure=configure)
    for m in (0, 2):
        z = np.mean(facs[m] == 0)
        assert 0.0 < z < 1.0
    assert abs(res["fit"] - fits[-1]) <= 1e-6 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-5)


@pytest.mark.parametrize("chunk", [50, 0])
def test_cpd_ntf_matches_oracle(gpu_available, chunk):
    # splatt_register_ntf_frob on mode 0 and splatt_register_ntf_lasso on
    # mode 2 (nonneg_frob.c, nonneg_lasso.c), chunked or one block
    lf, ll = 0.5, 1e-4

    def configure(L, co):
        m0 = np.array([0], dtype=np.uint64)
        m2 = np.array([2], dtype=np.uint64)
        assert L.splatt_register_ntf_frob(co, lf, m0.ctypes.data_as(ps.u64p), 1) == ps.SUCCESS
        assert L.splatt_register_ntf_lasso(co, ll, m2.ctypes.data_as(ps.u64p), 1) == ps.SUCCESS
        for m in range(3):
            co[0].chunk_sizes[m] = chunk
    res, facs, lam, fits = _run_both(tensor_path("med.tns"), 5, 3,
                                     cons=[(2, 8, 0.0, chunk, lf), (1, 0, 0.0),
                                           (2, 9, 0.0, chunk, ll)],
                                     unconstrained=False, configure=configure)
    for m in (0, 2):
        assert np.all(res["factors"][m] >= 0)
    assert abs(res["fit"] - fits[-1]) <= 1e-6 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-5)


@pytest.mark.parametrize("chunk", [50, 7])
def test_cpd_rowsimp_matches_oracle(gpu_available, chunk):
    # splatt_register_rowsimp on mode 0: chunked ADMM whose prox projects
    # every row onto the probability simplex
    def configure(L, co):
        modes = np.array([0], dtype=np.uint64)
        assert L.splatt_register_rowsimp(co, modes.ctypes.data_as(ps.u64p), 1) == ps.SUCCESS
        for m in range(3):
            co[0].chunk_sizes[m] = chunk
    res, facs, lam, fits = _run_both(tensor_path("med.tns"), 5, 3,
                                     cons=[(2, 5, 0.0, chunk), (1, 0, 0.0), (1, 0, 0.0)],
                                     unconstrained=False, configure=configure)
    A = res["factors"][0]
    assert np.all(A >= 0)
    assert np.allclose(A.sum(axis=1), 1.0, atol=1e-6)
    assert abs(res["fit"] - fits[-1]) <= 1e-6 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-5)


def test_cpd_smooth_matches_oracle(gpu_available):
    # splatt_register_smooth on mode 1: ADMM whose prox solves the banded
    # smoothness system down every column
    lam_s = 0.3

    def configure(L, co):
        modes = np.array([1], dtype=np.uint64)
        assert L.splatt_register_smooth(co, lam_s, modes.ctypes.data_as(ps.u64p), 1) == \
            ps.SUCCESS
    res, facs, lam, fits = _run_both(tensor_path("med.tns"), 5, 3,
                                     cons=[(1, 0, 0.0), (2, 6, 0.0, 0, lam_s), (1, 0, 0.0)],
                                     unconstrained=False, configure=configure)
    assert abs(res["fit"] - fits[-1]) <= 1e-6 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-5)


def test_cpd_colsimp_matches_oracle(gpu_available):
    # splatt_register_colsimp on mode 2: every column of the factor on the
    # probability simplex after each ADMM step
    def configure(L, co):
        modes = np.array([2], dtype=np.uint64)
        assert L.splatt_register_colsimp(co, modes.ctypes.data_as(ps.u64p), 1) == ps.SUCCESS
    res, facs, lam, fits = _run_both(tensor_path("med.tns"), 5, 3,
                                     cons=[(1, 0, 0.0), (1, 0, 0.0), (2, 7, 0.0)],
                                     unconstrained=False, configure=configure)
    C2 = res["factors"][2]
    assert np.all(C2 >= 0)
    assert np.allclose(C2.sum(axis=0), 1.0, atol=1e-6)
    assert abs(res["fit"] - fits[-1]) <= 1e-6 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-5)


def test_cpd_rowsimp_needs_chunks(gpu_available):
    def configure(L, co):
        modes = np.array([0], dtype=np.uint64)
        L.splatt_register_rowsimp(co, modes.ctypes.data_as(ps.u64p), 1)
        for m in range(3):
            co[0].chunk_sizes[m] = 0
    cs = _twomode(path=tensor_path("med.tns"))
    res = ps.run_cpd(cs, 5, max_its=1, tol=0.0, seed=7, configure=configure)
    assert res["rc"] == ps.ERROR_BADINPUT


@pytest.mark.parametrize("prox,lam", [(3, 0.0), (4, 2.0), (5, 0.0), (8, 0.5), (9, 2.0)])
def test_admm_chunked_matches_oracle(gpu_available, prox, lam):
    # the chunk kernel alone: rank 72 keeps the inverse in global memory,
    # FP32 MTTKRP input, 33-row chunks; nonneg, lasso, row simplex and the
    # non-negative frob and L1 forms
    I, R, chunk = 300, 72, 33
    rng = np.random.default_rng(31)
    A = rng.standard_normal((3 * R, R))
    G = A.T @ A + 0.5 * np.eye(R)
    M = (rng.standard_normal((I, R)) * 4.0).astype(np.float32)
    P0 = np.abs(rng.standard_normal((I, R)))
    D0 = rng.standard_normal((I, R)) * 0.1
    wantP, wantD, wtot = orc.admm_chunked(prox, M.astype(np.float64), P0, D0, G, chunk, 50,
                                          1e-2, lam=lam)
    L = ps.lib()
    RP = _rp(R)
    grams = [G, np.ones((R, R))]
    bufs = [_dev(g) for g in grams]
    arr = (C.c_void_p * 2)(*[b.ptr for b in bufs])
    ginv = ps.DevBuf(RP * RP * 8)
    rho = ps.DevBuf(8)
    info = ps.DevBuf(8)
    # two modes, mode 1 solved: the normal matrix is aTa[0] = G
    assert L.spd_dense_chol_inv(arr, 2, 1, R, 0.0, 1, ginv.ptr, rho.ptr, info.ptr) == 0
    dM, dP, dD = _dev(M), _dev(P0), _dev(D0)
    cnt = ps.DevBuf(8)
    assert L.spd_dense_admm_chunked(dM.ptr, 1, ginv.ptr, rho.ptr, I, R, chunk, dP.ptr, dD.ptr,
                                    prox, C.c_double(lam), 50, C.c_double(1e-2),
                                    cnt.ptr) == 0
    P = dP.to_host((I, R), np.float64)
    D = dD.to_host((I, R), np.float64)
    assert int(cnt.to_host((1,), np.uint64)[0]) == wtot
    tol = 1e-8
    assert np.allclose(P, wantP, rtol=tol, atol=tol * np.max(np.abs(wantP)))
    assert np.allclose(D, wantD, rtol=tol, atol=tol * np.max(np.abs(wantP)))


# ---- the benchmark workloads at their own kernel instantiations -----------

def test_cpd_uniform_1m_rank16_20_iterations(gpu_available):
    # CPD-ALS, rank 16, 20 iterations, FP64, 1000x800x600 with 1,000,000
    # uniform nonzeros (values U(0,1)), two-CSF allocation: the fit of every
    # iteration and the final model against the oracle from the same start
    print("case cpd 1000x800x600 1M R16", flush=True)
    inds, vals, dims = ps.synth_coord([1000, 800, 600], 1_000_000, vals="uniform",
                                      seed=20261203)
    cs = ps.CsfSet.convert(inds, vals, ps.Opts(csf_alloc=ps.CSF_TWOMODE))
    res = ps.run_cpd(cs, 16, max_its=20, tol=0.0, seed=1)
    assert res["rc"] == ps.SUCCESS, ps.last_error()
    coo = orc.Coo.from_arrays(inds, vals, dims)
    perm0 = cs.arrays(0)["perm"]
    ocsfs = [orc.Csf(coo, perm0), orc.Csf(coo, cs.arrays(1)["perm"])]
    mode_csf = [1 if perm0[-1] == m else 0 for m in range(3)]
    facs, lam, fits = orc.cpd_als_con(ocsfs, mode_csf, res["init"], 20, 0.0,
                                      [(0, 0, 0.0)] * 3, True, nthreads=8)
    assert len(fits) == 20
    assert abs(res["fit"] - fits[-1]) <= 1e-8 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-6)
    cs.free()


def _c3_tensor(nnz, seed=20261208):
    # the 4-mode workload's shape at a reduced nonzero count: Zipf(1.1) on
    # the first three modes, uniform on the fourth (2000), values U(-1,1),
    # compacted (no empty index in any mode: an untiled CSF stores no root
    # ids, so a root mode with empty slices would write slice positions)
    return ps.synth_coord([50_000, 30_000, 10_000, 2_000], nnz, zipf=[1.1, 1.1, 1.1, 0.0],
                          vals="symmetric", compact=True, seed=seed)


@pytest.mark.parametrize("prec", [0, 1])
def test_mttkrp_4mode_rank64_twomode(gpu_available, prec):
    # every mode of a two-CSF 4-mode tensor at rank 64 (the root, internal
    # and leaf walks at the rank-64 lane layout), FP64 and FP32, against the
    # long-double COO oracle
    print("case 4-mode rank 64 prec", prec, flush=True)
    inds, vals, _ = _c3_tensor(600_000)
    dims = [int(i.max()) + 1 for i in inds]
    coo = orc.Coo.from_arrays(inds, vals, dims)
    cs = ps.CsfSet.convert(inds, vals, ps.Opts(csf_alloc=ps.CSF_TWOMODE))
    R = 64
    mats = [ps.synth_matrix(d, R, 500 + m) for m, d in enumerate(dims)]
    L = ps.lib()
    assert L.splatt_amd_set_mttkrp_precision(prec) == ps.SUCCESS
    try:
        for mode in range(4):
            got = cs.mttkrp(mode, mats)
            if prec == 0:
                want = orc.mttkrp_coo(coo, mats, mode)
                tol = 1e-12
            else:
                m32 = [m.astype(np.float32) for m in mats]
                want = orc.mttkrp_coo(coo, m32, mode, f32=True)
                tol = 1e-5
            err = np.linalg.norm(got - want) / np.linalg.norm(want)
            assert err <= tol, (mode, err)
    finally:
        L.splatt_amd_set_mttkrp_precision(0)
    cs.free()


def test_cpd_4mode_rank64_fp64(gpu_available):
    # three FP64 CPD-ALS iterations of the 4-mode shape at rank 64 against
    # the oracle from the same start (the factors stay well posed over so
    # few iterations)
    print("case cpd 4-mode rank 64", flush=True)
    inds, vals, dims = _c3_tensor(300_000, seed=7)
    cs = ps.CsfSet.convert(inds, vals, ps.Opts(csf_alloc=ps.CSF_TWOMODE))
    res = ps.run_cpd(cs, 64, max_its=3, tol=0.0, seed=3)
    assert res["rc"] == ps.SUCCESS, ps.last_error()
    coo = orc.Coo.from_arrays(inds, vals, res["dims"])
    perm0 = cs.arrays(0)["perm"]
    ocsfs = [orc.Csf(coo, perm0), orc.Csf(coo, cs.arrays(1)["perm"])]
    mode_csf = [1 if perm0[-1] == m else 0 for m in range(4)]
    facs, lam, fits = orc.cpd_als_con(ocsfs, mode_csf, res["init"], 3, 0.0,
                                      [(0, 0, 0.0)] * 4, True, nthreads=8)
    assert abs(res["fit"] - fits[-1]) <= 1e-8 * max(1.0, abs(fits[-1]))
    _model_close(res, facs, lam, 1e-6)
    cs.free()
