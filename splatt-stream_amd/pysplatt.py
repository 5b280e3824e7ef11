"""ctypes binding of libsplatt.so (the SPLATT C API served by the GPU) and of
the synthetic tensor generator, for tests and benchmarks.

The binding mirrors the C structures field for field (include/splatt/
structs.h) so callers can read a CSF's arrays exactly as the reference's own
tests do. It raises if the native library is missing: there is no fallback.
"""
import ctypes as C
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsplatt.so")
GEN_PATH = os.path.join(HERE, "gen", "libsplatt_gen.so")

MAX_NMODES = 8
NOPTIONS = 12

# option indices, include/splatt/types.h
OPT_NTHREADS, OPT_TOLERANCE, OPT_REGULARIZE, OPT_NITER, OPT_VERBOSITY, \
    OPT_RANDSEED, OPT_CSF_ALLOC, OPT_TILE, OPT_TILELEVEL, OPT_PRIVTHRESH, \
    OPT_DECOMP, OPT_COMM = range(12)
CSF_ONEMODE, CSF_TWOMODE, CSF_ALLMODE = 0, 1, 2
NOTILE, DENSETILE = 0, 1
SUCCESS = 1
ERROR_BADINPUT = 2
F64, F32 = 0, 1
LAYOUT_AUTO, LAYOUT_CSF = 0, 1

u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


class CsfSparsity(C.Structure):
    _fields_ = [("nfibs", C.c_uint64 * MAX_NMODES),
                ("fptr", u64p * MAX_NMODES),
                ("fids", u64p * MAX_NMODES),
                ("vals", f64p)]


class SplattCsf(C.Structure):
    _fields_ = [("nnz", C.c_uint64),
                ("nmodes", C.c_uint64),
                ("dims", C.c_uint64 * MAX_NMODES),
                ("dim_perm", C.c_uint64 * MAX_NMODES),
                ("dim_iperm", C.c_uint64 * MAX_NMODES),
                ("which_tile", C.c_int),
                ("ntiles", C.c_uint64),
                ("ntiled_modes", C.c_uint64),
                ("tile_dims", C.c_uint64 * MAX_NMODES),
                ("pt", C.POINTER(CsfSparsity)),
                ("dev", C.c_void_p),
                ("alloc_count", C.c_uint64),
                ("alloc_type", C.c_int)]


class SplattCoord(C.Structure):
    _fields_ = [("nmodes", C.c_uint64),
                ("dims", C.c_uint64 * MAX_NMODES),
                ("nnz", C.c_uint64),
                ("ind", u64p * MAX_NMODES),
                ("vals", f64p),
                ("tiled", C.c_int),
                ("indmap", u64p * MAX_NMODES)]


class SplattKruskal(C.Structure):
    _fields_ = [("rank", C.c_uint64),
                ("factors", f64p * MAX_NMODES),
                ("lambda_", f64p),
                ("nmodes", C.c_uint64),
                ("dims", C.c_uint64 * MAX_NMODES),
                ("fit", C.c_double)]


class CpdOpts(C.Structure):
    _fields_ = [("tolerance", C.c_double),
                ("max_iterations", C.c_uint64),
                ("inner_tolerance", C.c_double),
                ("max_inner_iterations", C.c_uint64),
                ("constraints", C.c_void_p * MAX_NMODES),
                ("chunk_sizes", C.c_uint64 * MAX_NMODES)]


class CommInfo(C.Structure):
    _fields_ = [("world_rank", C.c_int), ("world_npes", C.c_int)]


class GlobalOpts(C.Structure):
    _fields_ = [("num_threads", C.c_int), ("verbosity", C.c_int),
                ("random_seed", C.c_int)]


class StreamResult(C.Structure):
    _fields_ = [("nmodes", C.c_uint64), ("rank", C.c_uint64),
                ("stream_mode", C.c_uint64), ("nbatches", C.c_uint64),
                ("dims", C.c_uint64 * MAX_NMODES),
                ("factors", f64p * MAX_NMODES),
                ("iperm", u64p * MAX_NMODES),
                ("local_err", f64p), ("batch_seconds", f64p),
                ("outer_its", u64p), ("final_err", C.c_double),
                ("global_err", f64p), ("local10_err", f64p), ("cpd_err", f64p)]


class MatrixT(C.Structure):
    _fields_ = [("I", C.c_uint64), ("J", C.c_uint64), ("vals", f64p),
                ("rowmajor", C.c_int)]


class MttkrpWs(C.Structure):
    _fields_ = [("num_csf", C.c_uint64),
                ("mode_csf_map", C.c_uint64 * MAX_NMODES),
                ("num_threads", C.c_uint64),
                ("tile_partition", u64p * MAX_NMODES),
                ("tree_partition", u64p * MAX_NMODES),
                ("is_privatized", C.c_bool * MAX_NMODES),
                ("privatize_buffer", C.c_void_p),
                ("reduction_time", C.c_double)]


_lib = None
_gen = None


def lib():
    """Load libsplatt.so (raises OSError if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libsplatt.so not built: run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.splatt_default_opts.restype = C.POINTER(C.c_double)
    L.splatt_free_opts.argtypes = [C.POINTER(C.c_double)]
    L.splatt_csf_convert.restype = C.c_int
    L.splatt_csf_convert.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(u64p), f64p,
                                     C.POINTER(C.POINTER(SplattCsf)),
                                     C.POINTER(C.c_double)]
    L.splatt_csf_load.restype = C.c_int
    L.splatt_csf_load.argtypes = [C.c_char_p, u64p, C.POINTER(C.POINTER(SplattCsf)),
                                  C.POINTER(C.c_double)]
    L.splatt_csf_alloc.restype = C.POINTER(SplattCsf)
    L.splatt_csf_alloc.argtypes = [C.POINTER(SplattCoord), C.POINTER(C.c_double)]
    L.splatt_free_csf.argtypes = [C.POINTER(SplattCsf), C.POINTER(C.c_double)]
    L.splatt_csf_frobsq.restype = C.c_double
    L.splatt_csf_frobsq.argtypes = [C.POINTER(SplattCsf)]
    L.splatt_csf_find_mode_order.argtypes = [u64p, C.c_uint64, C.c_int, C.c_uint64, u64p]
    L.splatt_mttkrp.restype = C.c_int
    L.splatt_mttkrp.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(SplattCsf),
                                C.POINTER(f64p), f64p, C.POINTER(C.c_double)]
    L.splatt_mttkrp_alloc_ws.restype = C.POINTER(MttkrpWs)
    L.splatt_mttkrp_alloc_ws.argtypes = [C.POINTER(SplattCsf), C.c_uint64,
                                         C.POINTER(C.c_double)]
    L.splatt_mttkrp_free_ws.argtypes = [C.POINTER(MttkrpWs)]
    L.splatt_tt_read.restype = C.POINTER(SplattCoord)
    L.splatt_tt_read.argtypes = [C.c_char_p]
    L.splatt_tt_free.argtypes = [C.POINTER(SplattCoord)]
    L.splatt_tt_remove_empty.restype = C.c_uint64
    L.splatt_tt_remove_empty.argtypes = [C.POINTER(SplattCoord)]
    L.splatt_alloc_coord.restype = C.POINTER(SplattCoord)
    L.splatt_alloc_coord.argtypes = [C.c_uint64, C.c_uint64]
    L.splatt_free_coord.argtypes = [C.POINTER(SplattCoord)]
    L.splatt_coord_load.restype = C.c_int
    L.splatt_coord_load.argtypes = [C.c_char_p, C.POINTER(SplattCoord)]
    L.splatt_amd_last_error.restype = C.c_char_p
    L.splatt_amd_set_mttkrp_precision.restype = C.c_int
    L.splatt_amd_set_mttkrp_precision.argtypes = [C.c_int]
    L.splatt_amd_set_mttkrp_layout.restype = C.c_int
    L.splatt_amd_set_mttkrp_layout.argtypes = [C.c_int]
    L.splatt_amd_mttkrp_dev.restype = C.c_int
    L.splatt_amd_mttkrp_dev.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(SplattCsf),
                                        C.POINTER(C.c_void_p), C.c_void_p, C.c_int,
                                        C.POINTER(C.c_double), C.c_void_p]
    L.splatt_amd_mttkrp_dev_noalloc.restype = C.c_int
    L.splatt_amd_mttkrp_dev_noalloc.argtypes = L.splatt_amd_mttkrp_dev.argtypes
    L.splatt_amd_mttkrp_noalloc_bytes.restype = C.c_size_t
    L.splatt_amd_mttkrp_noalloc_bytes.argtypes = []
    L.splatt_amd_mttkrp_cost.restype = C.c_int
    L.splatt_amd_mttkrp_cost.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(SplattCsf),
                                         C.c_int, C.POINTER(C.c_double),
                                         f64p, f64p]
    L.splatt_amd_mttkrp_timing.argtypes = [C.c_int]
    L.splatt_amd_mttkrp_last_ms.restype = C.c_double
    L.splatt_amd_mttkrp_kernel_times.restype = C.c_int
    L.splatt_amd_mttkrp_kernel_times.argtypes = [f64p, C.c_int]
    L.splatt_amd_dev_alloc.restype = C.c_void_p
    L.splatt_amd_dev_alloc.argtypes = [C.c_size_t]
    L.splatt_amd_dev_free.argtypes = [C.c_void_p]
    L.splatt_amd_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.splatt_amd_h2d.restype = C.c_int
    L.splatt_amd_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.splatt_amd_d2h.restype = C.c_int
    L.splatt_amd_sync.restype = C.c_int
    L.splatt_version_major.restype = C.c_int
    L.splatt_alloc_cpd.restype = C.POINTER(SplattKruskal)
    L.splatt_alloc_cpd.argtypes = [C.POINTER(SplattCsf), C.c_uint64]
    L.splatt_free_cpd.argtypes = [C.POINTER(SplattKruskal)]
    L.splatt_cpd.restype = C.c_int
    L.splatt_cpd.argtypes = [C.POINTER(SplattCsf), C.c_uint64, C.POINTER(CpdOpts),
                             C.POINTER(GlobalOpts), C.POINTER(SplattKruskal)]
    L.splatt_alloc_cpd_opts.restype = C.POINTER(CpdOpts)
    L.splatt_free_cpd_opts.argtypes = [C.POINTER(CpdOpts)]
    L.splatt_alloc_global_opts.restype = C.POINTER(GlobalOpts)
    L.splatt_free_global_opts.argtypes = [C.POINTER(GlobalOpts)]
    L.splatt_register_frob.restype = C.c_int
    L.splatt_register_frob.argtypes = [C.POINTER(CpdOpts), C.c_double, u64p, C.c_uint64]
    L.splatt_register_lasso.restype = C.c_int
    L.splatt_register_lasso.argtypes = [C.POINTER(CpdOpts), C.c_double, u64p, C.c_uint64]
    L.splatt_register_smooth.restype = C.c_int
    L.splatt_register_smooth.argtypes = [C.POINTER(CpdOpts), C.c_double, u64p, C.c_uint64]
    for name in ("splatt_register_ntf_frob", "splatt_register_ntf_lasso"):
        getattr(L, name).restype = C.c_int
        getattr(L, name).argtypes = [C.POINTER(CpdOpts), C.c_double, u64p, C.c_uint64]
    L.splatt_alloc_comm_info.restype = C.POINTER(CommInfo)
    L.splatt_alloc_comm_info.argtypes = []
    L.splatt_free_comm_info.argtypes = [C.POINTER(CommInfo)]
    for name in ("splatt_register_maxcolnorm", "splatt_register_maxcolnorm_nonneg",
                 "splatt_register_nonneg", "splatt_register_rowsimp",
                 "splatt_register_colsimp"):
        getattr(L, name).restype = C.c_int
        getattr(L, name).argtypes = [C.POINTER(CpdOpts), u64p, C.c_uint64]
    L.spd_dense_gram.restype = C.c_int
    L.spd_dense_gram.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                 C.c_void_p, C.c_void_p]
    L.spd_dense_chol_inv.restype = C.c_int
    L.spd_dense_chol_inv.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_uint64,
                                     C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
    L.spd_dense_apply.restype = C.c_int
    L.spd_dense_apply.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64,
                                  C.c_uint64, C.c_void_p, C.c_void_p]
    L.spd_dense_admm_step.restype = C.c_int
    L.spd_dense_admm_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    L.spd_dense_admm_solve.restype = C.c_int
    L.spd_dense_admm_solve.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                       C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                       C.c_uint64, C.c_double, C.c_void_p, C.c_void_p]
    L.spd_dense_admm_chunked.restype = C.c_int
    L.spd_dense_admm_chunked.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p,
                                         C.c_void_p, C.c_int, C.c_double, C.c_uint64,
                                         C.c_double, C.c_void_p]
    L.splatt_amd_cpd_stream_coo.restype = C.POINTER(StreamResult)
    L.splatt_amd_cpd_stream_coo.argtypes = [C.POINTER(SplattCoord), C.c_uint64, C.c_uint64,
                                            C.c_double, C.c_uint64, C.c_int,
                                            C.POINTER(CpdOpts), C.POINTER(GlobalOpts)]
    L.splatt_amd_free_stream_result.argtypes = [C.POINTER(StreamResult)]
    L.splatt_cpd_stream.restype = C.c_int
    L.splatt_cpd_stream.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64, C.c_double,
                                    C.POINTER(CpdOpts), C.POINTER(GlobalOpts)]
    L.splatt_amd_server_create.restype = C.c_void_p
    L.splatt_amd_coord_load_dev.argtypes = [C.c_char_p, C.POINTER(SplattCoord), u64p]
    L.splatt_amd_coord_load_dev.restype = C.c_int
    L.spd_decimal_convert.argtypes = [C.c_char_p, u64p, C.c_uint64, C.c_int, f64p, u64p,
                                      C.POINTER(C.c_int)]
    L.spd_decimal_convert.restype = C.c_int
    L.spd_coo_upload.argtypes = [C.c_int, C.c_uint64, u64p, C.POINTER(u64p), f64p,
                                 C.POINTER(C.c_void_p)]
    L.spd_coo_upload.restype = C.c_int
    L.spd_coo_free.argtypes = [C.c_void_p]
    L.spd_coo_free.restype = None
    L.spd_coo_remove_empty.argtypes = [C.c_void_p, u64p]
    L.spd_coo_remove_empty.restype = C.c_int
    L.spd_coo_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int), u64p, u64p]
    L.spd_coo_dims.restype = None
    L.spd_coo_download.argtypes = [C.c_void_p, C.POINTER(u64p), f64p]
    L.spd_coo_download.restype = C.c_int
    L.splatt_amd_server_create.argtypes = [C.c_size_t, C.c_int]
    L.splatt_amd_server_open.restype = C.c_int64
    L.splatt_amd_server_open.argtypes = [C.c_void_p, C.c_uint64, u64p, C.c_uint64, C.c_uint64,
                                         C.c_double, C.POINTER(CpdOpts), C.c_uint]
    L.splatt_amd_server_submit.restype = C.c_int64
    L.splatt_amd_server_submit.argtypes = [C.c_void_p, C.c_int64, C.c_uint64,
                                           C.POINTER(u64p), f64p]
    L.splatt_amd_server_drain.restype = C.c_int
    L.splatt_amd_server_drain.argtypes = [C.c_void_p]
    L.splatt_amd_server_slice.restype = C.c_int
    L.splatt_amd_server_slice.argtypes = [C.c_void_p, C.c_int64, f64p, f64p, f64p, u64p]
    L.splatt_amd_server_factors.restype = C.c_int
    L.splatt_amd_server_factors.argtypes = [C.c_void_p, C.c_int64, C.c_uint64, f64p, u64p]
    L.splatt_amd_server_close.restype = C.c_int
    L.splatt_amd_server_close.argtypes = [C.c_void_p, C.c_int64]
    L.splatt_amd_server_stats.argtypes = [C.c_void_p, C.POINTER(C.c_size_t),
                                          C.POINTER(C.c_size_t), C.POINTER(C.c_int),
                                          C.POINTER(C.c_int)]
    L.splatt_amd_server_free.argtypes = [C.c_void_p]
    _lib = L
    return L


def gen():
    global _gen
    if _gen is not None:
        return _gen
    if not os.path.exists(GEN_PATH):
        raise OSError("libsplatt_gen.so not built: run __graft_entry__.build()")
    G = C.CDLL(GEN_PATH)
    G.splatt_gen_coord.restype = C.c_int
    G.splatt_gen_coord.argtypes = [C.c_int, u64p, f64p, C.c_uint64, C.c_int, C.c_int,
                                   C.c_uint64, C.POINTER(u64p), f64p, u64p]
    G.splatt_gen_shuffle.argtypes = [C.c_int, C.c_uint64, C.POINTER(u64p), f64p,
                                     C.c_uint64]
    G.splatt_gen_matrix.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, f64p]
    G.splatt_gen_write_tns.argtypes = [C.c_char_p, C.c_int, C.c_uint64, C.POINTER(u64p), f64p]
    G.splatt_gen_write_tns.restype = C.c_int
    _gen = G
    return G


def last_error():
    return lib().splatt_amd_last_error().decode(errors="replace")


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Opts:
    """An options array from splatt_default_opts(), freed on drop."""

    def __init__(self, **kw):
        self._p = lib().splatt_default_opts()
        self.arr = np.ctypeslib.as_array(self._p, shape=(NOPTIONS,))
        for k, v in kw.items():
            self.arr[globals()["OPT_" + k.upper()]] = v

    @property
    def ptr(self):
        return self._p

    def __del__(self):
        if getattr(self, "_p", None) is not None and _lib is not None:
            _lib.splatt_free_opts(self._p)
            self._p = None


def synth_coord(dims, nnz, zipf=None, vals="uniform", compact=True, seed=1,
                shuffle=True):
    """Synthetic tensor: list of per-mode uint64 index arrays, float64 values
    and the (compacted) dims."""
    n = len(dims)
    zipf = [0.0] * n if zipf is None else list(zipf)
    kinds = {"uniform": 0, "lognormal": 1, "symmetric": 2}
    inds = [np.empty(nnz, dtype=np.uint64) for _ in range(n)]
    v = np.empty(nnz, dtype=np.float64)
    dims_in = _u64(dims)
    dims_out = np.zeros(n, dtype=np.uint64)
    zs = np.ascontiguousarray(zipf, dtype=np.float64)
    ptrs = (u64p * n)(*[_ptr(a, C.c_uint64) for a in inds])
    rc = gen().splatt_gen_coord(n, _ptr(dims_in, C.c_uint64), _ptr(zs, C.c_double),
                                nnz, kinds[vals], 1 if compact else 0, seed, ptrs,
                                _ptr(v, C.c_double), _ptr(dims_out, C.c_uint64))
    if rc != 0:
        raise ValueError("nnz exceeds the number of cells")
    if shuffle:
        gen().splatt_gen_shuffle(n, nnz, ptrs, _ptr(v, C.c_double), seed)
    return inds, v, [int(d) for d in dims_out]


def write_tns(path, inds, vals):
    """Write a COO as 1-indexed .tns text (values as %.17g)."""
    n = len(inds)
    keep = [_u64(a) for a in inds]
    v = np.ascontiguousarray(vals, dtype=np.float64)
    ptrs = (u64p * n)(*[_ptr(a, C.c_uint64) for a in keep])
    if gen().splatt_gen_write_tns(path.encode(), n, len(v), ptrs, _ptr(v, C.c_double)) != 0:
        raise OSError("writing " + path)


def synth_matrix(nrows, ncols, seed):
    out = np.empty((nrows, ncols), dtype=np.float64)
    gen().splatt_gen_matrix(nrows, ncols, seed, _ptr(out, C.c_double))
    return out


class CsfSet:
    """A CSF allocation (1, 2 or nmodes tensors) owned by libsplatt."""

    def __init__(self, ptr, opts):
        self.ptr = ptr
        self.opts = opts
        self.count = int(ptr[0].alloc_count)

    @classmethod
    def convert(cls, inds, vals, opts):
        n = len(inds)
        keep = [_u64(a) for a in inds]
        v = np.ascontiguousarray(vals, dtype=np.float64)
        ptrs = (u64p * n)(*[_ptr(a, C.c_uint64) for a in keep])
        out = C.POINTER(SplattCsf)()
        rc = lib().splatt_csf_convert(n, len(v), ptrs, _ptr(v, C.c_double),
                                      C.byref(out), opts.ptr)
        if rc != SUCCESS or not out:
            raise RuntimeError("splatt_csf_convert failed: " + last_error())
        return cls(out, opts)

    def free(self):
        if self.ptr:
            lib().splatt_free_csf(self.ptr, self.opts.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def tensor(self, i):
        return self.ptr[i]

    def arrays(self, i):
        """Copies of tensor i's host mirror: dict with perm, tiles (list of
        per-tile dicts of nfibs/fptr/fids/vals as numpy arrays or None)."""
        ct = self.ptr[i]
        N = int(ct.nmodes)
        tiles = []
        for t in range(int(ct.ntiles)):
            pt = ct.pt[t]
            nf = [int(pt.nfibs[l]) for l in range(N)]
            fptr, fids = [], []
            for l in range(N):
                if l < N - 1 and pt.fptr[l]:
                    fptr.append(np.ctypeslib.as_array(pt.fptr[l], shape=(nf[l] + 1,)).copy())
                else:
                    fptr.append(None)
                if pt.fids[l]:
                    fids.append(np.ctypeslib.as_array(pt.fids[l], shape=(max(nf[l], 1),))[:nf[l]].copy())
                else:
                    fids.append(None)
            vals = (np.ctypeslib.as_array(pt.vals, shape=(nf[N - 1],)).copy()
                    if pt.vals else None)
            tiles.append(dict(nfibs=nf, fptr=fptr, fids=fids, vals=vals))
        return dict(nnz=int(ct.nnz), nmodes=N, dims=[int(ct.dims[m]) for m in range(N)],
                    perm=[int(ct.dim_perm[m]) for m in range(N)],
                    iperm=[int(ct.dim_iperm[m]) for m in range(N)],
                    which_tile=int(ct.which_tile), ntiles=int(ct.ntiles),
                    ntiled_modes=int(ct.ntiled_modes),
                    tile_dims=[int(ct.tile_dims[m]) for m in range(N)], tiles=tiles)

    def mttkrp(self, mode, mats):
        """splatt_mttkrp with host factors; returns dims[mode] x R float64."""
        R = mats[0].shape[1]
        keep = [np.ascontiguousarray(m, dtype=np.float64) for m in mats]
        ptrs = (f64p * len(keep))(*[_ptr(m, C.c_double) for m in keep])
        out = np.empty((int(self.ptr[0].dims[mode]), R), dtype=np.float64)
        rc = lib().splatt_mttkrp(mode, R, self.ptr, ptrs, _ptr(out, C.c_double),
                                 self.opts.ptr)
        if rc != SUCCESS:
            raise RuntimeError("splatt_mttkrp failed: " + last_error())
        return out

    def frobsq(self, i=0):
        return lib().splatt_csf_frobsq(C.byref(self.ptr[i]))


class DevBuf:
    """A device allocation through libsplatt (no torch involved)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().splatt_amd_dev_alloc(self.nbytes)
        if not self.ptr:
            raise MemoryError("device allocation failed: " + last_error())

    @classmethod
    def from_host(cls, arr):
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        if lib().splatt_amd_h2d(b.ptr, a.ctypes.data, a.nbytes) != SUCCESS:
            raise RuntimeError("h2d failed: " + last_error())
        return b

    def to_host(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        if lib().splatt_amd_d2h(out.ctypes.data, self.ptr, out.nbytes) != SUCCESS:
            raise RuntimeError("d2h failed: " + last_error())
        return out

    def free(self):
        if self.ptr:
            lib().splatt_amd_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_libc = None


def srand(seed):
    """Seed glibc rand(), which splatt_alloc_cpd draws from."""
    global _libc
    if _libc is None:
        _libc = C.CDLL(None)
    _libc.srand(C.c_uint(seed))


def run_cpd(csfset, rank, max_its=20, tol=0.0, seed=1, configure=None,
            verbosity=0, max_inner=None, inner_tol=None):
    """splatt_alloc_cpd (after srand(seed)) + splatt_cpd through the C API.
    configure(L, opts_ptr) may register constraints. Returns a dict with the
    initial factors, the result factors, lambda, fit, the return code and
    cpd_s, the wall time of the splatt_cpd call alone."""
    L = lib()
    srand(seed)
    K = L.splatt_alloc_cpd(csfset.ptr, rank)
    N = int(K[0].nmodes)
    dims = [int(K[0].dims[m]) for m in range(N)]
    init = [np.ctypeslib.as_array(K[0].factors[m], shape=(dims[m], rank)).copy()
            for m in range(N)]
    co = L.splatt_alloc_cpd_opts()
    co[0].max_iterations = max_its
    co[0].tolerance = tol
    if max_inner is not None:
        co[0].max_inner_iterations = max_inner
    if inner_tol is not None:
        co[0].inner_tolerance = inner_tol
    if configure is not None:
        configure(L, co)
    go = L.splatt_alloc_global_opts()
    go[0].verbosity = verbosity
    t0 = time.perf_counter()
    rc = L.splatt_cpd(csfset.ptr, rank, co, go, K)
    cpd_s = time.perf_counter() - t0
    out = [np.ctypeslib.as_array(K[0].factors[m], shape=(dims[m], rank)).copy()
           for m in range(N)]
    lam = np.ctypeslib.as_array(K[0].lambda_, shape=(rank,)).copy()
    fit = float(K[0].fit)
    L.splatt_free_cpd_opts(co)
    L.splatt_free_global_opts(go)
    L.splatt_free_cpd(K)
    return dict(init=init, factors=out, lam=lam, fit=fit, rc=rc, dims=dims, cpd_s=cpd_s)


def coord_from_arrays(inds, vals, dims):
    """A splatt_coord owned by libsplatt holding copies of the arrays."""
    L = lib()
    n = len(inds)
    nnz = len(vals)
    tt = L.splatt_alloc_coord(n, nnz)
    for m in range(n):
        tt[0].dims[m] = int(dims[m])
        dst = np.ctypeslib.as_array(tt[0].ind[m], shape=(max(nnz, 1),))
        dst[:nnz] = np.asarray(inds[m], dtype=np.uint64)
    dv = np.ctypeslib.as_array(tt[0].vals, shape=(max(nnz, 1),))
    dv[:nnz] = np.asarray(vals, dtype=np.float64)
    return tt


def run_stream(inds, vals, dims, rank, stream_mode, forget=0.95, max_its=5, seed=1,
               frob=None, verbosity=0, check_interval=0, refit=False):
    """splatt_amd_cpd_stream_coo after srand(seed); frob registers Frobenius
    regularisation on the streamed mode; check_interval/refit drive the
    checkpoint errors (the reference uses 100 and a refit). Returns a dict of
    numpy results."""
    L = lib()
    tt = coord_from_arrays(inds, vals, dims)
    co = L.splatt_alloc_cpd_opts()
    co[0].max_iterations = max_its
    if frob is not None:
        modes = np.array([stream_mode], dtype=np.uint64)
        L.splatt_register_frob(co, frob, modes.ctypes.data_as(u64p), 1)
    go = L.splatt_alloc_global_opts()
    go[0].verbosity = verbosity
    srand(seed)
    r = L.splatt_amd_cpd_stream_coo(tt, rank, stream_mode, forget, check_interval,
                                    1 if refit else 0, co, go)
    out = None
    if r:
        R = r[0]
        N = int(R.nmodes)
        nb = int(R.nbatches)
        d = [int(R.dims[m]) for m in range(N)]
        facs = [np.ctypeslib.as_array(R.factors[m], shape=(max(d[m], 1), rank))[:d[m]].copy()
                for m in range(N)]
        iperm = [np.ctypeslib.as_array(R.iperm[m], shape=(max(int(dims[m]), 1),))[:int(dims[m])].copy()
                 for m in range(N)]
        out = dict(nbatches=nb, dims=d, factors=facs, iperm=iperm,
                   local_err=np.ctypeslib.as_array(R.local_err, shape=(max(nb, 1),))[:nb].copy(),
                   batch_seconds=np.ctypeslib.as_array(R.batch_seconds, shape=(max(nb, 1),))[:nb].copy(),
                   outer_its=np.ctypeslib.as_array(R.outer_its, shape=(max(nb, 1),))[:nb].copy(),
                   final_err=float(R.final_err),
                   global_err=np.ctypeslib.as_array(R.global_err, shape=(max(nb, 1),))[:nb].copy(),
                   local10_err=np.ctypeslib.as_array(R.local10_err, shape=(max(nb, 1),))[:nb].copy(),
                   cpd_err=np.ctypeslib.as_array(R.cpd_err, shape=(max(nb, 1),))[:nb].copy())
        L.splatt_amd_free_stream_result(r)
    L.splatt_free_cpd_opts(co)
    L.splatt_free_global_opts(go)
    L.splatt_free_coord(tt)
    return out


class Server:
    """The multi-stream server (splatt_amd_server_*)."""

    def __init__(self, budget=0, max_inflight=4):
        self.L = lib()
        self.ptr = self.L.splatt_amd_server_create(budget, max_inflight)
        if not self.ptr:
            raise RuntimeError("server create failed: " + last_error())
        self.ranks = {}
        self.nmodes = {}
        self.tmode = {}

    def open(self, dims_cap, rank, stream_mode, forget=0.95, max_its=5, frob=None, seed=1):
        co = self.L.splatt_alloc_cpd_opts()
        co[0].max_iterations = max_its
        if frob is not None:
            modes = np.array([stream_mode], dtype=np.uint64)
            self.L.splatt_register_frob(co, frob, modes.ctypes.data_as(u64p), 1)
        cap = np.ascontiguousarray(dims_cap, dtype=np.uint64)
        sid = self.L.splatt_amd_server_open(self.ptr, len(cap), cap.ctypes.data_as(u64p), rank,
                                            stream_mode, forget, co, seed)
        self.L.splatt_free_cpd_opts(co)
        if sid < 0:
            raise RuntimeError("server open failed: " + last_error())
        self.ranks[sid] = rank
        self.nmodes[sid] = len(cap)
        self.tmode[sid] = stream_mode
        return sid

    def submit(self, sid, inds, vals):
        keep = [np.ascontiguousarray(i, dtype=np.uint64) for i in inds]
        v = np.ascontiguousarray(vals, dtype=np.float64)
        ptrs = (u64p * len(keep))(*[k.ctypes.data_as(u64p) for k in keep])
        t = self.L.splatt_amd_server_submit(self.ptr, sid, len(v), ptrs, v.ctypes.data_as(f64p))
        if t < 0:
            raise RuntimeError("submit failed: " + last_error())
        return t

    def drain(self):
        if self.L.splatt_amd_server_drain(self.ptr) != SUCCESS:
            raise RuntimeError("drain failed: " + last_error())

    def slice(self, ticket):
        lat, gs, le = C.c_double(), C.c_double(), C.c_double()
        oi = C.c_uint64()
        rc = self.L.splatt_amd_server_slice(self.ptr, ticket, C.byref(lat), C.byref(gs),
                                            C.byref(le), C.byref(oi))
        if rc < 0:
            raise RuntimeError("slice failed: " + last_error())
        if rc == 1:
            return None
        return dict(latency=lat.value, gpu=gs.value, local_err=le.value, outer=oi.value)

    def factors(self, sid, mode):
        rows = C.c_uint64()
        self.L.splatt_amd_server_factors(self.ptr, sid, mode, None, C.byref(rows))
        out = np.zeros((max(rows.value, 1), self.ranks[sid]), dtype=np.float64)
        self.L.splatt_amd_server_factors(self.ptr, sid, mode, out.ctypes.data_as(f64p),
                                         C.byref(rows))
        return out[:rows.value]

    def close(self, sid):
        self.L.splatt_amd_server_close(self.ptr, sid)

    def stats(self):
        u, b = C.c_size_t(), C.c_size_t()
        a, w = C.c_int(), C.c_int()
        self.L.splatt_amd_server_stats(self.ptr, C.byref(u), C.byref(b), C.byref(a), C.byref(w))
        return dict(used=u.value, budget=b.value, admitted=a.value, waiting=w.value)

    def free(self):
        if self.ptr:
            self.L.splatt_amd_server_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
