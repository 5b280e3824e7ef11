"""GPU CSF construction: the host mirror built on the GPU equals the oracle's
CSF array for array (bit for bit), for every allocation scheme, with and
without dense tiling, on the reference fixtures and on synthetic Zipf
tensors; the caller's COO is left in CSF order as the reference leaves it.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import DATASETS, GOLDEN, tensor_path
import oracle_py as orc
import pysplatt as ps

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    assert a["perm"] == b["perm"], what
    assert a["ntiles"] == b["ntiles"], what
    assert a["tile_dims"] == b["tile_dims"], what
    assert a["ntiled_modes"] == b["ntiled_modes"], what
    N = a["nmodes"]
    for t, (x, y) in enumerate(zip(a["tiles"], b["tiles"])):
        assert x["nfibs"] == y["nfibs"], (what, t)
        for l in range(N):
            for key in ("fptr", "fids"):
                u, v = x[key][l], y[key][l]
                assert (u is None) == (v is None), (what, t, key, l)
                if u is not None:
                    assert np.array_equal(u, v), (what, t, key, l)
        assert (x["vals"] is None) == (y["vals"] is None)
        if x["vals"] is not None:
            assert np.array_equal(x["vals"], y["vals"]), (what, t)


def _oracle_set(coo, kind, tile=0, tilelevel=1, nthreads=1):
    dims = coo.dims
    N = len(dims)
    if kind == ps.CSF_ONEMODE:
        perms = [orc.mode_order(dims, 0)]
        tiles = [tile]
    elif kind == ps.CSF_TWOMODE:
        p0 = orc.mode_order(dims, 0)
        perms = [p0, orc.mode_order(dims, 3, p0[-1])]
        tiles = [tile, 0]
    else:
        perms = [orc.mode_order(dims, 3, m) for m in range(N)]
        tiles = [tile] * N
    return [orc.Csf(coo, p, tile=t, tilelevel=tilelevel, nthreads=nthreads).arrays()
            for p, t in zip(perms, tiles)]


@pytest.mark.parametrize("name", DATASETS)
@pytest.mark.parametrize("kind", [ps.CSF_ONEMODE, ps.CSF_TWOMODE, ps.CSF_ALLMODE])
def test_fixture_csf_bit_exact(gpu_available, name, kind):
    print("case", name, kind, flush=True)
    coo = orc.Coo.read(tensor_path(name))
    opts = ps.Opts(csf_alloc=kind)
    got = ps.CsfSet.convert(coo.inds(), coo.vals(), opts)
    want = _oracle_set(coo, kind)
    assert got.count == len(want)
    for i, w in enumerate(want):
        _same(got.arrays(i), w, (name, kind, i))


def test_small_golden(gpu_available):
    g = json.load(open(os.path.join(GOLDEN, "small_csf.json")))
    coo = orc.Coo.read(tensor_path("small.tns"))
    got = ps.CsfSet.convert(coo.inds(), coo.vals(), ps.Opts(csf_alloc=ps.CSF_ALLMODE))
    for i, key in ((0, "onemode"), (1, "minusone_1"), (2, "minusone_2")):
        a = got.arrays(i)
        t = a["tiles"][0]
        assert a["perm"] == g[key]["perm"]
        assert t["nfibs"] == g[key]["nfibs"]
        assert [list(x) for x in t["fptr"][:2]] == g[key]["fptr"]
        assert t["fids"][0] is None
        assert [list(x) for x in t["fids"][1:]] == g[key]["fids"][1:]
        assert list(t["vals"]) == g[key]["vals"]


@pytest.mark.parametrize("tilelevel,nthreads", [(1, 4), (2, 3), (3, 2)])
def test_densetile_bit_exact(gpu_available, tilelevel, nthreads):
    print("case densetile", tilelevel, nthreads, flush=True)
    for name in ("med.tns", "med4.tns"):
        coo = orc.Coo.read(tensor_path(name))
        opts = ps.Opts(csf_alloc=ps.CSF_TWOMODE, tile=ps.DENSETILE,
                       tilelevel=tilelevel, nthreads=nthreads)
        got = ps.CsfSet.convert(coo.inds(), coo.vals(), opts)
        want = _oracle_set(coo, ps.CSF_TWOMODE, tile=1, tilelevel=tilelevel,
                           nthreads=nthreads)
        for i, w in enumerate(want):
            _same(got.arrays(i), w, (name, tilelevel, nthreads, i))


@pytest.mark.parametrize("dims,nnz,zipf", [
    ([3000, 2000, 1000], 200000, [1.2, 1.2, 1.2]),
    ([500, 400, 300, 50], 150000, [1.1, 1.1, 1.1, 0.0]),
    ([70000, 60000, 50000], 100000, [0.0, 0.0, 0.0]),
    ([9, 8, 7, 6, 5, 4, 3, 2], 50000, [0.0] * 8),
])
def test_synthetic_bit_exact(gpu_available, dims, nnz, zipf):
    print("case synthetic", dims, nnz, flush=True)
    inds, vals, d2 = ps.synth_coord(dims, nnz, zipf=zipf, vals="lognormal", seed=21)
    coo = orc.Coo.from_arrays(inds, vals, d2)
    got = ps.CsfSet.convert(inds, vals, ps.Opts(csf_alloc=ps.CSF_ALLMODE))
    for i, w in enumerate(_oracle_set(coo, ps.CSF_ALLMODE)):
        _same(got.arrays(i), w, (dims, i))


def test_csf_alloc_reorders_coo_like_reference(gpu_available):
    # tests/csf_test.c:31-61 relies on csf_alloc leaving tt sorted
    L = ps.lib()
    p = L.splatt_tt_read(tensor_path("med4.tns").encode())
    opts = ps.Opts(csf_alloc=ps.CSF_ONEMODE, nthreads=1)
    cs = L.splatt_csf_alloc(p, opts.ptr)
    assert cs
    tt = p[0]
    nnz = int(tt.nnz)
    ct = cs[0]
    N = int(ct.nmodes)
    assert int(ct.pt[0].nfibs[0]) == int(tt.dims[int(ct.dim_perm[0])])
    leaf = np.ctypeslib.as_array(ct.pt[0].fids[N - 1], shape=(nnz,))
    assert np.array_equal(np.ctypeslib.as_array(tt.ind[int(ct.dim_perm[N - 1])], shape=(nnz,)), leaf)
    assert np.array_equal(np.ctypeslib.as_array(tt.vals, shape=(nnz,)),
                          np.ctypeslib.as_array(ct.pt[0].vals, shape=(nnz,)))
    # and the reordered COO is the oracle's reordered COO
    o = orc.Coo.read(tensor_path("med4.tns"))
    orc.Csf(o, orc.mode_order(o.dims, 0), reorder=True)
    for m in range(N):
        assert np.array_equal(np.ctypeslib.as_array(tt.ind[m], shape=(nnz,)), o.inds()[m])
    assert abs(L.splatt_csf_frobsq(cs) - float(np.sum(o.vals() ** 2))) <= 1e-9 * L.splatt_csf_frobsq(cs)
    L.splatt_free_csf(cs, opts.ptr)
    L.splatt_tt_free(p)


def test_csf_load(gpu_available):
    L = ps.lib()
    opts = ps.Opts(csf_alloc=ps.CSF_ONEMODE)
    out = ctypes.POINTER(ps.SplattCsf)()
    nm = ctypes.c_uint64(0)
    rc = L.splatt_csf_load(tensor_path("small4.tns").encode(), ctypes.byref(nm),
                           ctypes.byref(out), opts.ptr)
    assert rc == ps.SUCCESS and nm.value == 4
    o = orc.Coo.read(tensor_path("small4.tns"))
    o.remove_empty()
    want = orc.Csf(o, orc.mode_order(o.dims, 0)).arrays()
    s = ps.CsfSet(out, opts)
    _same(s.arrays(0), want, "csf_load")


def test_bad_input_is_rejected(gpu_available):
    # coordinates outside their dims fail with BADINPUT, not a fault
    inds = [np.array([0, 5], dtype=np.uint64), np.array([0, 1], dtype=np.uint64)]
    vals = np.array([1.0, 2.0])
    L = ps.lib()
    opts = ps.Opts(csf_alloc=ps.CSF_ONEMODE)
    p = L.splatt_alloc_coord(2, 2)
    p[0].dims[0] = 3
    p[0].dims[1] = 2
    for m in range(2):
        np.ctypeslib.as_array(p[0].ind[m], shape=(2,))[:] = inds[m]
    np.ctypeslib.as_array(p[0].vals, shape=(2,))[:] = vals
    cs = L.splatt_csf_alloc(p, opts.ptr)
    assert not cs
    L.splatt_free_coord(p)
