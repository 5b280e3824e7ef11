"""GPU MTTKRP parity against the oracle's brute-force COO MTTKRP (long
double accumulation, the reference's own known-answer procedure,
tests/mttkrp_test.c:36-88).

Tolerances (normwise relative, ||got - ref||_F / ||ref||_F):
  FP64 path 1e-12, FP32 path 1e-5  -- as stated for this path.
Covered: every mode of every allocation scheme (root, internal and leaf
kernels; internal and leaf outputs both over row-sorted copies and by the
CSF walk with atomics), 3- to 8-mode tensors, tiled tensors, ranks that exercise every
lane layout (1, 3, 8, 16, 32, 64, 100), Zipf-skewed tensors whose heavy
slices span many work items, the device-pointer entry point, and at a
larger size the mode-independence of <X, [[A_1..A_N]]> = sum(M_m * A_m).
"""
import ctypes

import numpy as np
import pytest

from conftest import DATASETS, tensor_path
import oracle_py as orc
import pysplatt as ps

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64))
                 / max(np.linalg.norm(np.asarray(b, np.float64)), 1e-300))


def _set_precision(p):
    assert ps.lib().splatt_amd_set_mttkrp_precision(p) == ps.SUCCESS


def _set_layout(x):
    assert ps.lib().splatt_amd_set_mttkrp_layout(x) == ps.SUCCESS


@pytest.fixture(autouse=True)
def _reset_precision():
    yield
    ps.lib().splatt_amd_set_mttkrp_precision(ps.F64)
    ps.lib().splatt_amd_set_mttkrp_layout(ps.LAYOUT_AUTO)


@pytest.mark.parametrize("name", DATASETS)
@pytest.mark.parametrize("kind", [ps.CSF_ONEMODE, ps.CSF_TWOMODE, ps.CSF_ALLMODE])
@pytest.mark.parametrize("layout", [ps.LAYOUT_AUTO, ps.LAYOUT_CSF])
def test_fixtures_fp64(gpu_available, name, kind, layout):
    print("case", name, kind, layout, flush=True)
    _set_layout(layout)
    coo = orc.Coo.read(tensor_path(name))
    s = ps.CsfSet.convert(coo.inds(), coo.vals(), ps.Opts(csf_alloc=kind))
    mats = [ps.synth_matrix(d, 16, 100 + m) for m, d in enumerate(coo.dims)]
    for mode in range(coo.nmodes):
        assert rel(s.mttkrp(mode, mats), orc.mttkrp_coo(coo, mats, mode)) <= TOL64


@pytest.mark.parametrize("name", ["med.tns", "med4.tns", "med5.tns"])
@pytest.mark.parametrize("layout", [ps.LAYOUT_AUTO, ps.LAYOUT_CSF])
def test_fixtures_fp32(gpu_available, name, layout):
    print("case fp32", name, layout, flush=True)
    _set_layout(layout)
    coo = orc.Coo.read(tensor_path(name))
    s = ps.CsfSet.convert(coo.inds(), coo.vals(), ps.Opts(csf_alloc=ps.CSF_ONEMODE))
    mats = [ps.synth_matrix(d, 32, 200 + m).astype(np.float32) for m, d in enumerate(coo.dims)]
    _set_precision(ps.F32)
    for mode in range(coo.nmodes):
        got = s.mttkrp(mode, [m.astype(np.float64) for m in mats])
        assert rel(got, orc.mttkrp_coo(coo, mats, mode, f32=True)) <= TOL32


@pytest.mark.parametrize("R", [1, 3, 8, 16, 32, 64, 100])
@pytest.mark.parametrize("prec", [ps.F64, ps.F32])
@pytest.mark.parametrize("layout", [ps.LAYOUT_AUTO, ps.LAYOUT_CSF])
def test_ranks_and_layouts(gpu_available, R, prec, layout):
    print("case rank", R, prec, layout, flush=True)
    _set_layout(layout)
    inds, vals, dims = ps.synth_coord([300, 200, 100], 30000, zipf=[1.2, 1.1, 0.0],
                                      vals="lognormal", seed=R)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    s = ps.CsfSet.convert(inds, vals, ps.Opts(csf_alloc=ps.CSF_ONEMODE))
    mats = [ps.synth_matrix(d, R, 7 + m) for m, d in enumerate(dims)]
    _set_precision(prec)
    for mode in range(3):
        got = s.mttkrp(mode, mats)
        if prec == ps.F64:
            assert rel(got, orc.mttkrp_coo(coo, mats, mode)) <= TOL64
        else:
            m32 = [m.astype(np.float32) for m in mats]
            assert rel(got, orc.mttkrp_coo(coo, m32, mode, f32=True)) <= TOL32


@pytest.mark.parametrize("nmodes", [4, 5, 6, 8])
@pytest.mark.parametrize("layout", [ps.LAYOUT_AUTO, ps.LAYOUT_CSF])
def test_higher_order(gpu_available, nmodes, layout):
    print("case nmodes", nmodes, layout, flush=True)
    _set_layout(layout)
    dims = [40, 30, 20, 12, 9, 7, 5, 4][:nmodes]
    inds, vals, d2 = ps.synth_coord(dims, 40000, zipf=[1.1] * nmodes, seed=nmodes)
    coo = orc.Coo.from_arrays(inds, vals, d2)
    mats = [ps.synth_matrix(d, 8, 30 + m) for m, d in enumerate(d2)]
    for kind in (ps.CSF_ONEMODE, ps.CSF_TWOMODE, ps.CSF_ALLMODE):
        s = ps.CsfSet.convert(inds, vals, ps.Opts(csf_alloc=kind))
        for mode in range(nmodes):
            assert rel(s.mttkrp(mode, mats), orc.mttkrp_coo(coo, mats, mode)) <= TOL64


@pytest.mark.parametrize("tilelevel", [1, 2])
def test_tiled(gpu_available, tilelevel):
    coo = orc.Coo.read(tensor_path("med4.tns"))
    opts = ps.Opts(csf_alloc=ps.CSF_TWOMODE, tile=ps.DENSETILE, tilelevel=tilelevel,
                   nthreads=5)
    s = ps.CsfSet.convert(coo.inds(), coo.vals(), opts)
    mats = [ps.synth_matrix(d, 16, 40 + m) for m, d in enumerate(coo.dims)]
    for mode in range(4):
        assert rel(s.mttkrp(mode, mats), orc.mttkrp_coo(coo, mats, mode)) <= TOL64


@pytest.mark.parametrize("layout", [ps.LAYOUT_AUTO, ps.LAYOUT_CSF])
def test_heavy_zipf_slices(gpu_available, layout):
    _set_layout(layout)
    # a few slices hold most nonzeros, so they span many work items and are
    # completed through the partial (atomic) flushes
    inds, vals, dims = ps.synth_coord([2000, 1500, 1000], 1000000, zipf=[1.6, 1.6, 1.6],
                                      vals="lognormal", seed=3)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    mats = [ps.synth_matrix(d, 32, 50 + m) for m, d in enumerate(dims)]
    for kind in (ps.CSF_ONEMODE, ps.CSF_ALLMODE):
        s = ps.CsfSet.convert(inds, vals, ps.Opts(csf_alloc=kind))
        for mode in range(3):
            assert rel(s.mttkrp(mode, mats), orc.mttkrp_coo(coo, mats, mode)) <= TOL64


def _edge_tensor(seed):
    """A 3-mode tensor whose mode-0 slices have lengths around the kernels'
    work-item (32 nonzeros), wave (256) and block (1024) edges, long slices
    spanning many waves next to runs of one-nonzero slices, and fibers of
    1 to 40 nonzeros: every path of the partial-slice combines (one slice per
    wave or block, several slices, empty items) and of the carry fix-up."""
    rng = np.random.default_rng(seed)
    pattern = [1, 31, 32, 33, 1, 1, 255, 256, 257, 2, 511, 512, 513, 1023, 1024, 1025,
               7, 1, 20000, 3, 1, 1, 1, 64, 96, 300, 1, 45000, 5, 129]
    flen = [1, 2, 7, 31, 33, 40, 3, 1, 1]
    D1, D2 = 4999, 3001
    i0, i1, i2 = [], [], []
    s = 0
    for rep in range(3):
        for L in pattern:
            f = 0
            left = L
            while left > 0:
                n = min(left, flen[(s + f) % len(flen)])
                j = (f * 37 + s) % D1
                ks = (np.arange(n) * 17 + f + s) % D2
                i0.append(np.full(n, s, np.uint64))
                i1.append(np.full(n, j, np.uint64))
                i2.append(ks.astype(np.uint64))
                left -= n
                f += 1
            s += 1
    inds = [np.concatenate(i0), np.concatenate(i1), np.concatenate(i2)]
    # no empty index in any mode (an untiled CSF's root rows are then its
    # slice positions, as in the reference)
    dims = [s]
    for m in (1, 2):
        used, inds[m] = np.unique(inds[m], return_inverse=True)
        inds[m] = inds[m].astype(np.uint64)
        dims.append(len(used))
    vals = rng.lognormal(0.0, 1.0, len(inds[0]))
    perm = rng.permutation(len(vals))
    return [a[perm] for a in inds], vals[perm], dims


@pytest.mark.parametrize("prec", [ps.F64, ps.F32])
@pytest.mark.parametrize("layout", [ps.LAYOUT_AUTO, ps.LAYOUT_CSF])
def test_slice_edges(gpu_available, prec, layout):
    print("case slice edges", prec, layout, flush=True)
    inds, vals, dims = _edge_tensor(11)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    mats = [ps.synth_matrix(d, 32, 70 + m) for m, d in enumerate(dims)]
    m32 = [m.astype(np.float32) for m in mats]
    _set_precision(prec)
    _set_layout(layout)
    for kind in (ps.CSF_ONEMODE, ps.CSF_TWOMODE, ps.CSF_ALLMODE):
        s = ps.CsfSet.convert(inds, vals, ps.Opts(csf_alloc=kind))
        try:
            for mode in range(3):
                got = s.mttkrp(mode, mats)
                if prec == ps.F64:
                    assert rel(got, orc.mttkrp_coo(coo, mats, mode)) <= TOL64, (kind, mode)
                else:
                    assert rel(got, orc.mttkrp_coo(coo, m32, mode, f32=True)) <= TOL32, (kind, mode)
        finally:
            s.free()


def test_slice_edges_two_level_fixup(gpu_available):
    # the carry fix-up's two-pass form (taken by long carry buffers, from
    # 131072 entries) forced on every size
    L = ps.lib()
    L.spd_mttkrp_set_fix_two_level.argtypes = [ctypes.c_uint64]
    assert L.spd_mttkrp_set_fix_two_level(0) == 0
    try:
        for prec in (ps.F64, ps.F32):
            test_slice_edges(gpu_available, prec, ps.LAYOUT_AUTO)
        _set_precision(ps.F64)
        test_heavy_zipf_slices(gpu_available, ps.LAYOUT_AUTO)
    finally:
        L.spd_mttkrp_set_fix_two_level(131072)
        _set_precision(ps.F64)
        _set_layout(ps.LAYOUT_AUTO)


def test_device_entry_point(gpu_available):
    inds, vals, dims = ps.synth_coord([500, 400, 300], 50000, zipf=[1.2] * 3, seed=9)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    opts = ps.Opts(csf_alloc=ps.CSF_ALLMODE)
    s = ps.CsfSet.convert(inds, vals, opts)
    R = 32
    mats = [ps.synth_matrix(d, R, 60 + m) for m, d in enumerate(dims)]
    L = ps.lib()
    for prec, dt in ((ps.F64, np.float64), (ps.F32, np.float32)):
        dm = [ps.DevBuf.from_host(m.astype(dt)) for m in mats]
        arr = (ctypes.c_void_p * 3)(*[b.ptr for b in dm])
        for mode in range(3):
            out = ps.DevBuf(dims[mode] * R * np.dtype(dt).itemsize)
            assert L.splatt_amd_mttkrp_dev(mode, R, s.ptr, arr, out.ptr, prec, opts.ptr,
                                           None) == ps.SUCCESS
            assert L.splatt_amd_sync() == ps.SUCCESS
            got = out.to_host((dims[mode], R), dt)
            if prec == ps.F64:
                assert rel(got, orc.mttkrp_coo(coo, mats, mode)) <= TOL64
            else:
                m32 = [m.astype(np.float32) for m in mats]
                assert rel(got, orc.mttkrp_coo(coo, m32, mode, f32=True)) <= TOL32


def test_inner_product_is_mode_independent(gpu_available):
    # <X, [[A_1,...,A_N]]> = sum_i sum_r M_m[i,r] A_m[i,r] for every mode m
    inds, vals, dims = ps.synth_coord([20000, 15000, 5000], 5000000, zipf=[1.2] * 3,
                                      vals="lognormal", seed=12)
    s = ps.CsfSet.convert(inds, vals, ps.Opts(csf_alloc=ps.CSF_ONEMODE))
    mats = [ps.synth_matrix(d, 16, 70 + m) for m, d in enumerate(dims)]
    inner = [float(np.sum(s.mttkrp(m, mats) * mats[m])) for m in range(3)]
    scale = float(np.sum(np.abs(vals))) * 1.0
    assert max(inner) - min(inner) <= 1e-10 * scale
    # linearity in the values
    s2 = ps.CsfSet.convert(inds, 2.0 * vals, ps.Opts(csf_alloc=ps.CSF_ONEMODE))
    assert rel(s2.mttkrp(0, mats), 2.0 * s.mttkrp(0, mats)) <= 1e-14


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return h


def test_device_entry_point_two_caller_streams(gpu_available):
    # splatt_amd_mttkrp_dev returns without synchronising and accepts any
    # stream: back-to-back calls on two caller streams (no wait between
    # them) must not share partial-row buffers. Both results against the
    # oracle, for every mode of an all-mode CSF (each a root output with
    # carried partial rows), FP32 and FP64.
    inds, vals, dims = ps.synth_coord([3000, 2500, 2000], 400000, zipf=[1.2] * 3, seed=91)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    opts = ps.Opts(csf_alloc=ps.CSF_ALLMODE)
    s = ps.CsfSet.convert(inds, vals, opts)
    R = 32
    sets = [[ps.synth_matrix(d, R, 300 + 10 * k + m) for m, d in enumerate(dims)]
            for k in range(2)]
    L = ps.lib()
    h = _hip()
    streams = [ctypes.c_void_p(), ctypes.c_void_p()]
    for st in streams:
        assert h.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        for prec, dt in ((ps.F64, np.float64), (ps.F32, np.float32)):
            dms = [[ps.DevBuf.from_host(m.astype(dt)) for m in mats] for mats in sets]
            arrs = [(ctypes.c_void_p * 3)(*[b.ptr for b in dm]) for dm in dms]
            for mode in range(3):
                outs = [ps.DevBuf(dims[mode] * R * np.dtype(dt).itemsize) for _ in range(2)]
                for k in range(2):
                    assert L.splatt_amd_mttkrp_dev(mode, R, s.ptr, arrs[k], outs[k].ptr, prec,
                                                   opts.ptr, streams[k]) == ps.SUCCESS
                for st in streams:
                    assert h.hipStreamSynchronize(st) == 0
                for k in range(2):
                    got = outs[k].to_host((dims[mode], R), dt)
                    if prec == ps.F64:
                        assert rel(got, orc.mttkrp_coo(coo, sets[k], mode)) <= TOL64
                    else:
                        m32 = [m.astype(np.float32) for m in sets[k]]
                        assert rel(got, orc.mttkrp_coo(coo, m32, mode, f32=True)) <= TOL32
    finally:
        for st in streams:
            h.hipStreamDestroy(st)


@pytest.mark.parametrize("R", [16, 32, 48, 64, 100])
@pytest.mark.parametrize("prec", [ps.F64, ps.F32])
def test_root_kernel_variants(gpu_available, R, prec):
    # every 3-mode root kernel variant against the oracle and against the
    # default (per-nonzero k_root3): 1 the persistent kernel with hot leaf
    # rows in LDS (one A gather per fiber, runs carried across a wave's
    # chunks), 2 the same with every row from HBM, 3 k_root3 with one A
    # gather per fiber. Untiled all-mode CSF (root rows are slice positions)
    # and a tiled one (root ids stored). Zipf(1.5) roots give slices that
    # span many waves; at R=64 FP64 only 64 hot rows fit in LDS, so
    # hotter-ranked rows beyond them are read from HBM through the rank table.
    print("case root variants", R, prec, flush=True)
    inds, vals, dims = ps.synth_coord([4000, 3000, 2000], 600000, zipf=[1.5, 1.3, 1.2],
                                      vals="lognormal", seed=17 + R)
    coo = orc.Coo.from_arrays(inds, vals, dims)
    mats = [ps.synth_matrix(d, R, 900 + m) for m, d in enumerate(dims)]
    m32 = [m.astype(np.float32) for m in mats]
    L = ps.lib()
    _set_precision(prec)
    for opts in (ps.Opts(csf_alloc=ps.CSF_ALLMODE),
                 ps.Opts(csf_alloc=ps.CSF_ALLMODE, tile=ps.DENSETILE, tilelevel=1, nthreads=4)):
        s = ps.CsfSet.convert(inds, vals, opts)
        try:
            for mode in range(3):
                assert L.spd_mttkrp_set_root_variant(0) == 0
                base = s.mttkrp(mode, mats)
                for v in (1, 2, 3):
                    assert L.spd_mttkrp_set_root_variant(v) == 0
                    got = s.mttkrp(mode, mats)
                    if prec == ps.F64:
                        assert rel(got, orc.mttkrp_coo(coo, mats, mode)) <= TOL64, (mode, v)
                        assert rel(got, base) <= TOL64, (mode, v)
                    else:
                        assert rel(got, orc.mttkrp_coo(coo, m32, mode, f32=True)) <= TOL32, (mode, v)
                        assert rel(got, base) <= TOL32, (mode, v)
        finally:
            L.spd_mttkrp_set_root_variant(0)
            s.free()
