"""ctypes wrapper of liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / CPU baseline. Never part of
the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAXM = 8

u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)


class OrCoo(C.Structure):
    _fields_ = [("nmodes", C.c_uint64), ("nnz", C.c_uint64),
                ("dims", C.c_uint64 * MAXM), ("ind", u64p * MAXM), ("vals", f64p)]


class OrTile(C.Structure):
    _fields_ = [("nfibs", C.c_uint64 * MAXM), ("fptr", u64p * MAXM),
                ("fids", u64p * MAXM), ("vals", f64p)]


class OrCsf(C.Structure):
    _fields_ = [("nnz", C.c_uint64), ("nmodes", C.c_uint64),
                ("dims", C.c_uint64 * MAXM), ("dim_perm", C.c_uint64 * MAXM),
                ("dim_iperm", C.c_uint64 * MAXM), ("which_tile", C.c_int),
                ("ntiles", C.c_uint64), ("ntiled_modes", C.c_uint64),
                ("tile_dims", C.c_uint64 * MAXM), ("pt", C.POINTER(OrTile))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        L.or_tns_read.restype = C.POINTER(OrCoo)
        L.or_tns_read.argtypes = [C.c_char_p]
        L.or_coo_alloc.restype = C.POINTER(OrCoo)
        L.or_coo_alloc.argtypes = [C.c_uint64, C.c_uint64]
        L.or_coo_free.argtypes = [C.POINTER(OrCoo)]
        L.or_remove_empty.restype = C.c_uint64
        L.or_remove_empty.argtypes = [C.POINTER(OrCoo)]
        L.or_mode_order.argtypes = [u64p, C.c_uint64, C.c_int, C.c_uint64, u64p]
        L.or_csf_build.restype = C.POINTER(OrCsf)
        L.or_csf_build.argtypes = [C.POINTER(OrCoo), u64p, C.c_int, C.c_uint64,
                                   C.c_uint64, C.c_int]
        L.or_csf_free.argtypes = [C.POINTER(OrCsf)]
        L.or_csf_frobsq.restype = C.c_double
        L.or_csf_frobsq.argtypes = [C.POINTER(OrCsf)]
        L.or_mttkrp_coo.argtypes = [C.POINTER(OrCoo), C.POINTER(f64p), C.c_uint64,
                                    C.c_uint64, f64p]
        L.or_mttkrp_coo_f.argtypes = [C.POINTER(OrCoo), C.POINTER(f32p), C.c_uint64,
                                      C.c_uint64, f32p]
        L.or_mttkrp_csf.argtypes = [C.POINTER(OrCsf), C.POINTER(f64p), C.c_uint64,
                                    C.c_uint64, f64p, C.c_int]
        L.or_mttkrp_csf_f.argtypes = [C.POINTER(OrCsf), C.POINTER(f32p), C.c_uint64,
                                      C.c_uint64, f32p, C.c_int]
        L.or_fill_rand.argtypes = [f64p, C.c_uint64]
        L.or_cpd_als.restype = C.c_int
        L.or_cpd_als.argtypes = [C.POINTER(C.POINTER(OrCsf)), u64p, C.c_uint64,
                                 C.c_uint64, C.POINTER(f64p), f64p, C.c_uint64,
                                 C.c_double, f64p, C.c_int]
        L.or_mat_aTa.argtypes = [f64p, C.c_uint64, C.c_uint64, f64p]
        L.or_cholesky.restype = C.c_int
        L.or_cholesky.argtypes = [f64p, C.c_uint64]
        L.or_cholesky_solve.argtypes = [f64p, C.c_uint64, f64p, C.c_uint64]
        L.or_mat_normalize.argtypes = [f64p, C.c_uint64, C.c_uint64, f64p]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n):
    """Threads of the oracle's row-independent dense loops (results do not
    depend on the count)."""
    lib().or_set_threads(int(n))


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Coo:
    """An or_coo owned by the oracle library."""

    def __init__(self, ptr):
        self.ptr = ptr

    @classmethod
    def from_arrays(cls, inds, vals, dims):
        n = len(inds)
        nnz = len(vals)
        p = lib().or_coo_alloc(n, nnz)
        for m in range(n):
            p[0].dims[m] = int(dims[m])
            dst = np.ctypeslib.as_array(p[0].ind[m], shape=(max(nnz, 1),))
            dst[:nnz] = np.asarray(inds[m], dtype=np.uint64)
        dv = np.ctypeslib.as_array(p[0].vals, shape=(max(nnz, 1),))
        dv[:nnz] = np.asarray(vals, dtype=np.float64)
        return cls(p)

    @classmethod
    def read(cls, path):
        p = lib().or_tns_read(path.encode())
        if not p:
            raise IOError(path)
        return cls(p)

    @property
    def nmodes(self):
        return int(self.ptr[0].nmodes)

    @property
    def nnz(self):
        return int(self.ptr[0].nnz)

    @property
    def dims(self):
        return [int(self.ptr[0].dims[m]) for m in range(self.nmodes)]

    def inds(self):
        return [np.ctypeslib.as_array(self.ptr[0].ind[m], shape=(max(self.nnz, 1),))[:self.nnz].copy()
                for m in range(self.nmodes)]

    def vals(self):
        return np.ctypeslib.as_array(self.ptr[0].vals, shape=(max(self.nnz, 1),))[:self.nnz].copy()

    def remove_empty(self):
        return int(lib().or_remove_empty(self.ptr))

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.or_coo_free(self.ptr)
            self.ptr = None


def mode_order(dims, typ, mode=0):
    d = np.asarray(dims, dtype=np.uint64)
    perm = np.zeros(len(dims), dtype=np.uint64)
    lib().or_mode_order(_p(d, C.c_uint64), len(dims), typ, mode, _p(perm, C.c_uint64))
    return [int(x) for x in perm]


class Csf:
    def __init__(self, coo, perm, tile=0, tilelevel=1, nthreads=1, reorder=False):
        p = np.asarray(perm, dtype=np.uint64)
        self.ptr = lib().or_csf_build(coo.ptr, _p(p, C.c_uint64), tile, tilelevel,
                                      nthreads, 1 if reorder else 0)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.or_csf_free(self.ptr)
            self.ptr = None

    def arrays(self):
        ct = self.ptr[0]
        N = int(ct.nmodes)
        tiles = []
        for t in range(int(ct.ntiles)):
            pt = ct.pt[t]
            nf = [int(pt.nfibs[l]) for l in range(N)]
            fptr, fids = [], []
            for l in range(N):
                fptr.append(np.ctypeslib.as_array(pt.fptr[l], shape=(nf[l] + 1,)).copy()
                            if (l < N - 1 and pt.fptr[l]) else None)
                fids.append(np.ctypeslib.as_array(pt.fids[l], shape=(max(nf[l], 1),))[:nf[l]].copy()
                            if pt.fids[l] else None)
            vals = (np.ctypeslib.as_array(pt.vals, shape=(nf[N - 1],)).copy()
                    if pt.vals else None)
            tiles.append(dict(nfibs=nf, fptr=fptr, fids=fids, vals=vals))
        return dict(nnz=int(ct.nnz), nmodes=N, dims=[int(ct.dims[m]) for m in range(N)],
                    perm=[int(ct.dim_perm[m]) for m in range(N)],
                    iperm=[int(ct.dim_iperm[m]) for m in range(N)],
                    which_tile=int(ct.which_tile), ntiles=int(ct.ntiles),
                    ntiled_modes=int(ct.ntiled_modes),
                    tile_dims=[int(ct.tile_dims[m]) for m in range(N)], tiles=tiles)

    def frobsq(self):
        return lib().or_csf_frobsq(self.ptr)

    def mttkrp(self, mats, mode, nthreads=1, f32=False):
        t = np.float32 if f32 else np.float64
        ct = C.c_float if f32 else C.c_double
        keep = [np.ascontiguousarray(m, dtype=t) for m in mats]
        R = keep[0].shape[1]
        ptrs = (C.POINTER(ct) * len(keep))(*[_p(m, ct) for m in keep])
        out = np.empty((int(self.ptr[0].dims[mode]), R), dtype=t)
        fn = lib().or_mttkrp_csf_f if f32 else lib().or_mttkrp_csf
        fn(self.ptr, ptrs, R, mode, _p(out, ct), nthreads)
        return out


def mttkrp_coo(coo, mats, mode, f32=False):
    t = np.float32 if f32 else np.float64
    ct = C.c_float if f32 else C.c_double
    keep = [np.ascontiguousarray(m, dtype=t) for m in mats]
    R = keep[0].shape[1]
    ptrs = (C.POINTER(ct) * len(keep))(*[_p(m, ct) for m in keep])
    out = np.empty((coo.dims[mode], R), dtype=t)
    fn = lib().or_mttkrp_coo_f if f32 else lib().or_mttkrp_coo
    fn(coo.ptr, ptrs, R, mode, _p(out, ct))
    return out


def fill_rand(n):
    out = np.empty(n, dtype=np.float64)
    lib().or_fill_rand(_p(out, C.c_double), n)
    return out


def cpd_als(csfs, mode_csf, factors, max_its, tol, nthreads=1):
    """Runs the oracle CPD-ALS in place on copies of `factors`; returns
    (factors, lambda, fits)."""
    N = len(factors)
    R = factors[0].shape[1]
    facs = [np.ascontiguousarray(f, dtype=np.float64).copy() for f in factors]
    lam = np.zeros(R, dtype=np.float64)
    fits = np.zeros(max_its, dtype=np.float64)
    cptrs = (C.POINTER(OrCsf) * len(csfs))(*[c.ptr for c in csfs])
    mc = np.asarray(mode_csf, dtype=np.uint64)
    fptrs = (f64p * N)(*[_p(f, C.c_double) for f in facs])
    its = lib().or_cpd_als(cptrs, _p(mc, C.c_uint64), N, R, fptrs, _p(lam, C.c_double),
                           max_its, tol, _p(fits, C.c_double), nthreads)
    return facs, lam, fits[:its]


class _OrCon(C.Structure):
    _fields_ = [("kind", C.c_int), ("prox", C.c_int), ("diag_add", C.c_double),
                ("chunk", C.c_uint64), ("lam", C.c_double)]


def cpd_als_con(csfs, mode_csf, factors, max_its, tol, cons, unconstrained,
                max_inner=50, inner_tol=1e-2, nthreads=1):
    """or_cpd_als_con on copies of `factors`. cons: list of (kind, prox,
    diag_add[, chunk[, lam]]) per mode (lam: the lasso multiplier). Returns
    (factors, lambda, fits)."""
    N = len(factors)
    R = factors[0].shape[1]
    facs = [np.ascontiguousarray(f, dtype=np.float64).copy() for f in factors]
    lam = np.zeros(R, dtype=np.float64)
    fits = np.zeros(max_its, dtype=np.float64)
    cptrs = (C.POINTER(OrCsf) * len(csfs))(*[c.ptr for c in csfs])
    mc = np.asarray(mode_csf, dtype=np.uint64)
    fptrs = (f64p * N)(*[_p(f, C.c_double) for f in facs])
    carr = (_OrCon * 8)()
    for m, c in enumerate(cons):
        carr[m].kind, carr[m].prox, carr[m].diag_add = c[0], c[1], c[2]
        carr[m].chunk = c[3] if len(c) > 3 else 0
        carr[m].lam = c[4] if len(c) > 4 else 0.0
    L = lib()
    L.or_cpd_als_con.restype = C.c_int
    its = L.or_cpd_als_con(cptrs, _p(mc, C.c_uint64), C.c_uint64(N), C.c_uint64(R),
                           fptrs, _p(lam, C.c_double), C.c_uint64(max_its),
                           C.c_double(tol), C.c_uint64(max_inner), C.c_double(inner_tol),
                           carr, C.c_int(1 if unconstrained else 0),
                           _p(fits, C.c_double), C.c_int(nthreads))
    return facs, lam, fits[:its]


def admm(prox, M, P, D, G, max_inner, inner_tol, lam=0.0):
    """or_admm on copies; returns (P, D, iterations, rho)."""
    M = np.ascontiguousarray(M, dtype=np.float64)
    P = np.ascontiguousarray(P, dtype=np.float64).copy()
    D = np.ascontiguousarray(D, dtype=np.float64).copy()
    G = np.ascontiguousarray(G, dtype=np.float64).copy()
    I, R = P.shape
    rho = C.c_double(0.)
    L = lib()
    L.or_admm.restype = C.c_uint64
    its = L.or_admm(C.c_int(prox), _p(M, C.c_double), _p(P, C.c_double),
                    _p(D, C.c_double), _p(G, C.c_double), C.c_uint64(I), C.c_uint64(R),
                    C.c_uint64(max_inner), C.c_double(inner_tol), C.byref(rho),
                    C.c_double(lam))
    return P, D, int(its), rho.value


def admm_chunked(prox, M, P, D, G, chunk, max_inner, inner_tol, lam=0.0):
    """or_admm_chunked on copies; returns (P, D, iterations x rows summed)."""
    M = np.ascontiguousarray(M, dtype=np.float64)
    P = np.ascontiguousarray(P, dtype=np.float64).copy()
    D = np.ascontiguousarray(D, dtype=np.float64).copy()
    G = np.ascontiguousarray(G, dtype=np.float64).copy()
    I, R = P.shape
    L = lib()
    L.or_admm_chunked.restype = C.c_uint64
    tot = L.or_admm_chunked(C.c_int(prox), _p(M, C.c_double), _p(P, C.c_double),
                            _p(D, C.c_double), _p(G, C.c_double), C.c_uint64(I),
                            C.c_uint64(R), C.c_uint64(chunk), C.c_uint64(max_inner),
                            C.c_double(inner_tol), C.c_double(lam))
    return P, D, int(tot)


def prox(kind, P, thr=0.0):
    P = np.ascontiguousarray(P, dtype=np.float64).copy()
    I, R = P.shape
    lib().or_prox(C.c_int(kind), _p(P, C.c_double), C.c_uint64(I), C.c_uint64(R),
                  C.c_double(thr))
    return P


def smooth_prox(P, lam, rho):
    P = np.ascontiguousarray(P, dtype=np.float64).copy()
    I, R = P.shape
    lib().or_smooth_prox(_p(P, C.c_double), C.c_uint64(I), C.c_uint64(R), C.c_double(lam),
                         C.c_double(rho))
    return P


# ---- streaming ---------------------------------------------------------------

class _OrStreamCfg(C.Structure):
    _fields_ = [("stream_mode", C.c_uint64), ("rank", C.c_uint64),
                ("forget", C.c_double), ("max_its", C.c_uint64),
                ("cons", _OrCon * MAXM), ("check_interval", C.c_uint64),
                ("refit", C.c_int)]


class _OrStreamOut(C.Structure):
    _fields_ = [("max_batches", C.c_uint64), ("time", f64p),
                ("factors", f64p * MAXM), ("local_err", f64p),
                ("outer_its", u64p), ("nbatches", C.c_uint64),
                ("dims", C.c_uint64 * MAXM), ("final_err", C.c_double),
                ("global_err", f64p), ("local10_err", f64p), ("cpd_err", f64p)]


def mttkrp_stream(coo, mats, mode):
    R = mats[0].shape[1]
    keep = [np.ascontiguousarray(m, dtype=np.float64) for m in mats]
    ptrs = (f64p * len(keep))(*[_p(m, C.c_double) for m in keep])
    out = np.zeros((coo.dims[mode], R), dtype=np.float64)
    lib().or_mttkrp_stream(coo.ptr, ptrs, C.c_uint64(R), C.c_uint64(mode), _p(out, C.c_double))
    return out


def cpd_error(coo, factors, lam):
    R = factors[0].shape[1]
    keep = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    ptrs = (f64p * len(keep))(*[_p(f, C.c_double) for f in keep])
    fd = np.array([f.shape[0] for f in keep], dtype=np.uint64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    L = lib()
    L.or_cpd_error.restype = C.c_double
    return L.or_cpd_error(coo.ptr, ptrs, _p(fd, C.c_uint64), C.c_uint64(R), _p(lam, C.c_double))


def stream_prepare(coo, stream_mode):
    """or_stream_prepare in place; returns iperm per mode (new -> original)."""
    N = coo.nmodes
    ip = [np.zeros(max(d, 1), dtype=np.uint64) for d in coo.dims]
    arr = (u64p * MAXM)(*([_p(a, C.c_uint64) for a in ip] + [None] * (MAXM - N)))
    lib().or_stream_prepare(coo.ptr, C.c_uint64(stream_mode), arr)
    return [a[:d] for a, d in zip(ip, coo.dims)]


def stream_cpd(coo, stream_mode, rank, forget, max_its, cons, check_interval=0,
               refit=False):
    """or_stream_cpd on a prepared tensor (caller seeds rand()). cons: list of
    (kind, prox, diag_add) per mode; check_interval/refit drive the
    checkpoint errors. Returns a dict of numpy results."""
    N = coo.nmodes
    dims = coo.dims
    T = dims[stream_mode]
    cfg = _OrStreamCfg()
    cfg.stream_mode, cfg.rank, cfg.forget, cfg.max_its = stream_mode, rank, forget, max_its
    for m, (k, p, d) in enumerate(cons):
        cfg.cons[m].kind, cfg.cons[m].prox, cfg.cons[m].diag_add = k, p, d
    cfg.check_interval = check_interval
    cfg.refit = 1 if refit else 0
    time = np.zeros((max(T, 1), rank), dtype=np.float64)
    facs = [np.zeros((max(d, 1), rank), dtype=np.float64) for d in dims]
    lerr = np.zeros(max(T, 1), dtype=np.float64)
    oits = np.zeros(max(T, 1), dtype=np.uint64)
    out = _OrStreamOut()
    out.max_batches = T
    out.time = _p(time, C.c_double)
    for m in range(N):
        if m != stream_mode:
            out.factors[m] = _p(facs[m], C.c_double)
    out.local_err = _p(lerr, C.c_double)
    out.outer_its = _p(oits, C.c_uint64)
    chk = [np.full(max(T, 1), -1.0) for _ in range(3)]
    out.global_err, out.local10_err, out.cpd_err = [_p(a, C.c_double) for a in chk]
    lib().or_stream_cpd(coo.ptr, C.byref(cfg), C.byref(out))
    nb = int(out.nbatches)
    fd = [int(out.dims[m]) for m in range(N)]
    return dict(time=time[:nb], factors=[f[:fd[m]] for m, f in enumerate(facs)],
                local_err=lerr[:nb], outer_its=oits[:nb], final_err=float(out.final_err),
                nbatches=nb, dims=fd, global_err=chk[0][:nb], local10_err=chk[1][:nb],
                cpd_err=chk[2][:nb])
