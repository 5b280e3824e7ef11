"""The allocation-free MTTKRP entry point (no counterpart in the reference):
splatt_amd_mttkrp_dev_noalloc against splatt_amd_mttkrp_dev through the
C ABI, on two calls in a row, the second with a smaller input (its partial
rows then sit in a fixed array that still holds the first call's).

Bar: the path's tolerance (FP64 1e-12, FP32 1e-5, normwise relative): both
entry points run the same kernels, but slices that span many waves are
completed with floating-point atomics whose order varies between launches.
Also: a CSF that has not been through splatt_amd_mttkrp_dev is rejected with
an error, not computed.
"""
import ctypes

import numpy as np
import pytest

import pysplatt as ps

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _run(L, fn, s, opts, mode, R, dims, dm, prec, dt):
    arr = (ctypes.c_void_p * len(dm))(*[b.ptr for b in dm])
    out = ps.DevBuf(dims[mode] * R * np.dtype(dt).itemsize)
    try:
        rc = fn(mode, R, s.ptr, arr, out.ptr, prec, opts.ptr, None)
        assert L.splatt_amd_sync() == ps.SUCCESS
        return rc, (out.to_host((dims[mode], R), dt) if rc == ps.SUCCESS else None)
    finally:
        out.free()


def _case(dims, nnz, zipf, seed, kind, R, prec, dt, tol):
    L = ps.lib()
    inds, vals, dims = ps.synth_coord(dims, nnz, zipf=zipf, vals="lognormal", seed=seed)
    opts = ps.Opts(csf_alloc=kind)
    s = ps.CsfSet.convert(inds, vals, opts)
    mats = [ps.synth_matrix(d, R, seed + 10 + m).astype(dt) for m, d in enumerate(dims)]
    dm = [ps.DevBuf.from_host(m) for m in mats]
    try:
        for mode in range(len(dims)):
            rc, want = _run(L, L.splatt_amd_mttkrp_dev, s, opts, mode, R, dims, dm, prec, dt)
            assert rc == ps.SUCCESS, ps.last_error()
            rc, got = _run(L, L.splatt_amd_mttkrp_dev_noalloc, s, opts, mode, R, dims, dm,
                           prec, dt)
            assert rc == ps.SUCCESS, ps.last_error()
            assert _rel(got, want) <= tol, (mode, _rel(got, want))
    finally:
        for b in dm:
            b.free()
        s.free()


@pytest.mark.parametrize("prec,dt,tol", [(ps.F32, np.float32, 1e-5), (ps.F64, np.float64, 1e-12)])
def test_noalloc_equals_allocating_entry(gpu_available, prec, dt, tol):
    print("case noalloc all-mode", prec, flush=True)
    # all-mode CSF: every output is a root output on the persistent kernel;
    # the second, smaller tensor reuses the array the first one filled
    _case([20000, 15000, 5000], 2000000, [1.2] * 3, 5, ps.CSF_ALLMODE, 32, prec, dt, tol)
    _case([3000, 2000, 1000], 100000, [1.2] * 3, 6, ps.CSF_ALLMODE, 32, prec, dt, tol)
    print("case noalloc one-mode", prec, flush=True)
    # one CSF: the root, then internal and leaf outputs (row-sorted copies)
    _case([3000, 2000, 1000], 200000, [1.2] * 3, 7, ps.CSF_ONEMODE, 16, prec, dt, tol)
    _case([600, 500, 400], 30000, [1.2] * 3, 8, ps.CSF_ONEMODE, 16, prec, dt, tol)


def test_noalloc_rejects_unprepared_csf(gpu_available):
    print("case noalloc unprepared", flush=True)
    L = ps.lib()
    inds, vals, dims = ps.synth_coord([3000, 2000, 1000], 100000, zipf=[1.2] * 3, seed=9)
    opts = ps.Opts(csf_alloc=ps.CSF_ALLMODE)
    s = ps.CsfSet.convert(inds, vals, opts)
    R = 32
    mats = [ps.synth_matrix(d, R, 40 + m) for m, d in enumerate(dims)]
    dm = [ps.DevBuf.from_host(m) for m in mats]
    try:
        rc, _ = _run(L, L.splatt_amd_mttkrp_dev_noalloc, s, opts, 0, R, dims, dm, ps.F64,
                     np.float64)
        assert rc != ps.SUCCESS
        assert "allocation-free" in ps.last_error()
        assert L.splatt_amd_mttkrp_noalloc_bytes() == 16 << 20
    finally:
        for b in dm:
            b.free()
        s.free()
