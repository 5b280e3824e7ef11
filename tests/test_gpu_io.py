"""The text tensor reader on the GPU (splatt_amd_coord_load_dev, and the
input of splatt_csf_load): coordinates, values and dims equal the host
reader's and the oracle's (src/io.c:62-108, :273-348) bit for bit, on the
reference's fixtures, on lines of every shape the format allows (comments,
blank and one-character lines, tabs, CR LF, a last line without '\\n', fields
only the C library converts), on a large file, and when so many lines need
the C library that the whole file is read on the host. Empty-index removal
on the device equals the host's (src/sptensor.c:190-252), and CSFs loaded
from a file equal those built from the host reader's tensor."""
import ctypes
import os

import numpy as np
import pytest

from conftest import DATASETS, tensor_path
import oracle_py as orc
import pysplatt as ps

pytestmark = pytest.mark.gpu


def _coord_arrays(p):
    tt = p[0]
    N, nnz = int(tt.nmodes), int(tt.nnz)
    inds = [np.ctypeslib.as_array(tt.ind[m], shape=(max(nnz, 1),))[:nnz].copy()
            for m in range(N)]
    vals = np.ctypeslib.as_array(tt.vals, shape=(max(nnz, 1),))[:nnz].copy()
    return dict(nmodes=N, nnz=nnz, dims=[int(tt.dims[m]) for m in range(N)], inds=inds,
                vals=vals)


def _load_dev(path):
    """(rc, arrays or None, lines converted on the host)."""
    L = ps.lib()
    c = L.splatt_alloc_coord(1, 1)
    nh = ctypes.c_uint64(0)
    rc = L.splatt_amd_coord_load_dev(path.encode(), c, ctypes.byref(nh))
    out = _coord_arrays(c) if rc == ps.SUCCESS else None
    L.splatt_free_coord(c)
    return rc, out, int(nh.value)


def _load_host(path):
    L = ps.lib()
    p = L.splatt_tt_read(path.encode())
    if not p:
        return None
    out = _coord_arrays(p)
    L.splatt_tt_free(p)
    return out


def _same_coord(a, b, what):
    assert a["nmodes"] == b["nmodes"] and a["nnz"] == b["nnz"], what
    assert a["dims"] == b["dims"], what
    for m in range(a["nmodes"]):
        assert np.array_equal(a["inds"][m], b["inds"][m]), (what, m)
    # bit for bit, signed zeros included
    assert np.array_equal(a["vals"].view(np.uint64), b["vals"].view(np.uint64)), what


def _oracle(path):
    o = orc.Coo.read(path)
    return dict(nmodes=o.nmodes, nnz=o.nnz, dims=o.dims, inds=o.inds(), vals=o.vals())


def _write(path, text):
    with open(path, "wb") as f:
        f.write(text.encode() if isinstance(text, str) else text)
    return str(path)


@pytest.mark.parametrize("name", DATASETS + ["small4_zeroidx.tns"])
def test_fixtures_match_host_and_oracle(gpu_available, name):
    path = tensor_path(name)
    print("fixture", name, flush=True)
    rc, dev, nh = _load_dev(path)
    assert rc == ps.SUCCESS, ps.last_error()
    _same_coord(dev, _load_host(path), name + " host")
    _same_coord(dev, _oracle(path), name + " oracle")
    assert nh == 0          # every field of the fixtures is converted on the GPU


_ODD = {
    # comments, blank lines, one-character lines, last line without '\n'
    "comments": "# c\n\n1 2 3 1.5\n#mid\n \n2 1 1 -2.25e-3\n\n3 3 2 7\n4 1 1 0.1",
    "zeroidx": "0 1 1 1.0\n2 1 3 2.0\n1 2 2 3.0\n5",
    "trailing_blank": "1 2 3 4.5 \n2 3 4 5.5 \n",
    "tabs_crlf": "1\t2\t1\t1e2\r\n2   1   2   -0.000001\r\n3 3 3 +.5\r\n",
    # fields only the C library converts: the line goes to the host
    "host_lines": ("1 1 1 0x1p3\n2 2 2 inf\n3 3 3 -nan\n4 4 4 1e400\n5 5 5 4.9e-324\n"
                   "6 6 6 12345678901234567890.5\n7 7 7 1.5abc\n8 8 8\n9 9\n"
                   "10 10 10 1,5\n+11 012 13 .25\n12 12 12 1e-310\n13 13 13 -0\n"
                   "14 14 14 1e\n15 15 15 \x001\n16 16a 16 2\n17 17 17 2.2250738585072011e-308\n"),
    # signed zeros, exponent forms, the ends of the normal range
    "values": ("1 1 1 -0.0\n2 2 2 0e5\n3 3 3 1.7976931348623157e308\n"
               "4 4 4 2.2250738585072014e-308\n5 5 5 9007199254740993\n6 6 6 1E-5\n"),
    "onefield_after_blank": "\n\n\n1 1 1 1\n",
    # smallest index 2: read as 1-indexed (the smallest index is taken
    # against a start of 1, so the 0-or-1 check of src/io.c never fails)
    "twoidx": "2 2 2 1.0\n3 2 2 1.0\n",
}


@pytest.mark.parametrize("case", sorted(_ODD))
def test_odd_lines_match_host(gpu_available, tmp_path, case):
    path = _write(tmp_path / (case + ".tns"), _ODD[case])
    print("case", case, flush=True)
    rc, dev, nh = _load_dev(path)
    host = _load_host(path)
    assert rc == ps.SUCCESS, ps.last_error()
    _same_coord(dev, host, case)
    if case == "host_lines":
        # all but "+11 012 13 .25" and "-0"
        assert nh == 15
        assert _oracle(path)["nnz"] == dev["nnz"]
    else:
        _same_coord(dev, _oracle(path), case + " oracle")
        # lines the GPU conversion leaves to the C library: in "comments" the
        # line of one blank (data with no digits); in "trailing_blank" both
        # lines, whose trailing blank makes a fourth index field "4.5"
        assert nh == {"comments": 1, "trailing_blank": 2}.get(case, 0)


def test_bad_files_are_rejected(gpu_available, tmp_path):
    L = ps.lib()
    # only comments, one field per line, a file that is not there
    for name, text in (("comments", "# a\n#b\n"), ("onefield", "5\n6\n")):
        path = _write(tmp_path / (name + ".tns"), text)
        assert _load_host(path) is None, name
        rc, _, _ = _load_dev(path)
        assert rc == ps.ERROR_BADINPUT, name
        opts = ps.Opts(csf_alloc=ps.CSF_ONEMODE)
        out = ctypes.POINTER(ps.SplattCsf)()
        nm = ctypes.c_uint64(0)
        assert L.splatt_csf_load(path.encode(), ctypes.byref(nm), ctypes.byref(out),
                                 opts.ptr) == ps.ERROR_BADINPUT
    rc, _, _ = _load_dev(str(tmp_path / "absent.tns"))
    assert rc == ps.ERROR_BADINPUT


def _random_text(n, nmodes, seed, one_indexed=True):
    rng = np.random.default_rng(seed)
    ind = rng.integers(1 if one_indexed else 0, 40000, size=(n, nmodes))
    val = rng.standard_normal(n) * 10.0 ** rng.integers(-30, 30, size=n)
    fmts = ["%.17g", "%g", "%.6e", "%.3f", "%.0f", "%d"]
    which = rng.integers(0, len(fmts), size=n)
    seps = [" ", "\t", "  "]
    lines = []
    for k in range(n):
        f = fmts[which[k]]
        v = val[k] if f != "%d" else int(rng.integers(-1000, 1000))
        sep = seps[k % 3]
        lines.append(sep.join(str(int(x)) for x in ind[k]) + sep + (f % v))
        if k % 1013 == 0:
            lines.append("# comment %d" % k)
        if k % 2039 == 0:
            lines.append("")
    return "\n".join(lines) + ("\n" if seed % 2 else "")


@pytest.mark.parametrize("nmodes,seed", [(3, 1), (4, 2), (5, 3), (8, 4)])
def test_large_random_text_matches_host(gpu_available, tmp_path, nmodes, seed):
    n = 600_000 if nmodes <= 4 else 200_000
    path = _write(tmp_path / "big.tns", _random_text(n, nmodes, seed))
    print("random text", nmodes, "modes", n, "lines", flush=True)
    rc, dev, nh = _load_dev(path)
    assert rc == ps.SUCCESS, ps.last_error()
    _same_coord(dev, _load_host(path), "random")
    assert dev["nnz"] == n
    # %.3f and %.0f of the largest values carry more than 19 digits
    assert nh < n // 10


def test_many_host_lines_read_on_host(gpu_available, tmp_path):
    # more lines need the C library than the GPU path hands back: the file
    # is read on the host instead, with the same result
    n = (1 << 20) + 5000
    rng = np.random.default_rng(9)
    ind = rng.integers(1, 1000, size=(n, 3))
    lines = ["%d %d %d %s" % (a, b, c, float(v).hex())
             for (a, b, c), v in zip(ind, rng.standard_normal(n))]
    path = _write(tmp_path / "hex.tns", "\n".join(lines) + "\n")
    print("hex lines", n, flush=True)
    rc, dev, nh = _load_dev(path)
    assert rc == ps.SUCCESS, ps.last_error()
    assert nh == n
    _same_coord(dev, _load_host(path), "hex")


def _upload(inds, vals, dims):
    L = ps.lib()
    n = len(inds)
    keep = [np.ascontiguousarray(a, dtype=np.uint64) for a in inds]
    v = np.ascontiguousarray(vals, dtype=np.float64)
    d = np.ascontiguousarray(dims, dtype=np.uint64)
    ptrs = (ps.u64p * n)(*[a.ctypes.data_as(ps.u64p) for a in keep])
    h = ctypes.c_void_p()
    rc = L.spd_coo_upload(n, len(v), d.ctypes.data_as(ps.u64p), ptrs,
                          v.ctypes.data_as(ps.f64p), ctypes.byref(h))
    assert rc == 0, ps.last_error()
    return h


@pytest.mark.parametrize("dims,nnz", [([3000, 200, 5000], 4000), ([50, 60, 70, 80], 300),
                                      ([100000, 7, 40000], 200000)])
def test_remove_empty_matches_host(gpu_available, dims, nnz):
    inds, vals, _ = ps.synth_coord(dims, nnz, zipf=[1.1] * len(dims), compact=False, seed=4)
    L = ps.lib()
    h = _upload(inds, vals, dims)
    removed = ctypes.c_uint64(0)
    assert L.spd_coo_remove_empty(h, ctypes.byref(removed)) == 0, ps.last_error()
    nm = ctypes.c_int(0)
    n = ctypes.c_uint64(0)
    dd = np.zeros(8, dtype=np.uint64)
    L.spd_coo_dims(h, ctypes.byref(nm), ctypes.byref(n), dd.ctypes.data_as(ps.u64p))
    got = [np.empty(nnz, dtype=np.uint64) for _ in dims]
    gv = np.empty(nnz, dtype=np.float64)
    ptrs = (ps.u64p * len(dims))(*[a.ctypes.data_as(ps.u64p) for a in got])
    assert L.spd_coo_download(h, ptrs, gv.ctypes.data_as(ps.f64p)) == 0
    L.spd_coo_free(h)
    # the host's compaction of the same tensor
    p = L.splatt_alloc_coord(len(dims), nnz)
    for m in range(len(dims)):
        p[0].dims[m] = dims[m]
        np.ctypeslib.as_array(p[0].ind[m], shape=(nnz,))[:] = inds[m]
    np.ctypeslib.as_array(p[0].vals, shape=(nnz,))[:] = vals
    want_removed = L.splatt_tt_remove_empty(p)
    want = _coord_arrays(p)
    L.splatt_free_coord(p)
    assert removed.value == want_removed > 0
    assert [int(x) for x in dd[:len(dims)]] == want["dims"]
    for m in range(len(dims)):
        assert np.array_equal(got[m], want["inds"][m]), m
    assert np.array_equal(gv, vals)


def _csf_arrays(s):
    return [s.arrays(i) for i in range(s.count)]


@pytest.mark.parametrize("name", ["med.tns", "med4.tns", "med5.tns", "small4_zeroidx.tns"])
@pytest.mark.parametrize("kind", ["ONEMODE", "TWOMODE", "ALLMODE"])
def test_csf_load_equals_host_built(gpu_available, name, kind):
    # splatt_csf_load (file -> GPU -> compaction -> CSF) against the host
    # reader's tensor compacted on the host and given to splatt_csf_alloc
    L = ps.lib()
    path = tensor_path(name)
    opts = ps.Opts(csf_alloc=getattr(ps, "CSF_" + kind))
    out = ctypes.POINTER(ps.SplattCsf)()
    nm = ctypes.c_uint64(0)
    assert L.splatt_csf_load(path.encode(), ctypes.byref(nm), ctypes.byref(out),
                             opts.ptr) == ps.SUCCESS, ps.last_error()
    got = ps.CsfSet(out, opts)
    p = L.splatt_tt_read(path.encode())
    L.splatt_tt_remove_empty(p)
    opts2 = ps.Opts(csf_alloc=getattr(ps, "CSF_" + kind))
    cs = L.splatt_csf_alloc(p, opts2.ptr)
    assert cs
    want = ps.CsfSet(cs, opts2)
    a, b = _csf_arrays(got), _csf_arrays(want)
    assert len(a) == len(b) and int(nm.value) == a[0]["nmodes"]
    for i, (x, y) in enumerate(zip(a, b)):
        assert x["dims"] == y["dims"] and x["perm"] == y["perm"], (name, kind, i)
        for t, (u, v) in enumerate(zip(x["tiles"], y["tiles"])):
            assert u["nfibs"] == v["nfibs"], (name, kind, i, t)
            for l in range(x["nmodes"]):
                for key in ("fptr", "fids"):
                    assert (u[key][l] is None) == (v[key][l] is None)
                    if u[key][l] is not None:
                        assert np.array_equal(u[key][l], v[key][l]), (name, kind, i, key, l)
            assert np.array_equal(u["vals"], v["vals"]), (name, kind, i, t)
    got.free()
    want.free()
    L.splatt_tt_free(p)
