"""Host-side checks that need no GPU: the C-ABI library loads and exports
everything include/ declares, option defaults, file input, mode ordering,
compaction and the MTTKRP workspace map (all host C code)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from conftest import DATASETS, ROOT, tensor_path
import oracle_py as orc
import pysplatt as ps


def _declared_functions():
    names = set()
    pat = re.compile(r"^[A-Za-z_][\w \*]*?\b(splatt_\w+)\s*\(", re.M)
    for h in glob.glob(os.path.join(ROOT, "include", "**", "*.h"), recursive=True):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in pat.finditer(text):
            names.add(m.group(1))
    return sorted(names)


def test_exports_every_declared_symbol():
    L = ps.lib()
    names = _declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def test_version():
    assert ps.lib().splatt_version_major() == 2


def test_default_opts():
    o = ps.Opts()
    a = o.arr
    assert a[ps.OPT_TOLERANCE] == 1e-5
    assert a[ps.OPT_NITER] == 50
    assert a[ps.OPT_CSF_ALLOC] == ps.CSF_TWOMODE
    assert a[ps.OPT_TILE] == ps.NOTILE
    assert a[ps.OPT_TILELEVEL] == 1
    assert a[ps.OPT_PRIVTHRESH] == 0.02
    assert a[ps.OPT_NTHREADS] >= 1


@pytest.mark.parametrize("name", DATASETS + ["small4_zeroidx.tns"])
def test_tt_read_matches_oracle(name):
    L = ps.lib()
    p = L.splatt_tt_read(tensor_path(name).encode())
    assert p
    tt = p[0]
    o = orc.Coo.read(tensor_path(name))
    N = int(tt.nmodes)
    nnz = int(tt.nnz)
    assert N == o.nmodes and nnz == o.nnz
    assert [int(tt.dims[m]) for m in range(N)] == o.dims
    for m in range(N):
        assert np.array_equal(np.ctypeslib.as_array(tt.ind[m], shape=(nnz,)), o.inds()[m])
    assert np.array_equal(np.ctypeslib.as_array(tt.vals, shape=(nnz,)), o.vals())
    L.splatt_tt_free(p)


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return str(path)


_EDGE_FILES = {
    # comments, blank lines, a final line without '\n'
    "comments": "# a comment\n\n1 2 3 1.5\n#mid\n2 1 1 -2.25e-3\n\n3 3 2 7\n4 1 1 0.1",
    # a 0-indexed mode and a one-character last line (skipped, as getline
    # returns 1 for it)
    "zeroidx": "0 1 1 1.0\n2 1 3 2.0\n1 2 2 3.0\n5",
    # a trailing blank on the first data line is one more token
    "trailing": "1 2 3 4.5 \n2 3 4 5.5 \n",
    # tabs as separators, several spaces, a value in exponent form
    "tabs": "1\t2\t1\t1e2\n2   1   2   -0.000001\n",
}


@pytest.mark.parametrize("case", sorted(_EDGE_FILES))
def test_tt_read_edge_cases_match_oracle(tmp_path, case):
    path = _write(tmp_path / (case + ".tns"), _EDGE_FILES[case])
    L = ps.lib()
    p = L.splatt_tt_read(path.encode())
    assert p
    tt = p[0]
    o = orc.Coo.read(path)
    N, nnz = int(tt.nmodes), int(tt.nnz)
    assert N == o.nmodes and nnz == o.nnz
    assert [int(tt.dims[m]) for m in range(N)] == o.dims
    for m in range(N):
        assert np.array_equal(np.ctypeslib.as_array(tt.ind[m], shape=(nnz,)), o.inds()[m])
    assert np.array_equal(np.ctypeslib.as_array(tt.vals, shape=(nnz,)), o.vals())
    L.splatt_tt_free(p)


def test_tt_read_many_lines_matches_oracle(tmp_path):
    # enough lines that every reader thread gets a range, with comments and
    # blank lines at range edges; values printed with 17 significant digits
    rng = np.random.default_rng(3)
    n = 200_000
    ind = rng.integers(1, 5000, size=(n, 3))
    val = rng.standard_normal(n)
    lines = ["%d %d %d %.17g" % (a, b, c, v) for (a, b, c), v in zip(ind, val)]
    for k in range(0, n, 997):
        lines[k] = "# skipped\n\n" + lines[k]
    path = _write(tmp_path / "many.tns", "\n".join(lines) + "\n")
    L = ps.lib()
    p = L.splatt_tt_read(path.encode())
    tt = p[0]
    o = orc.Coo.read(path)
    nnz = int(tt.nnz)
    assert nnz == o.nnz == n
    for m in range(3):
        assert np.array_equal(np.ctypeslib.as_array(tt.ind[m], shape=(nnz,)), o.inds()[m])
    assert np.array_equal(np.ctypeslib.as_array(tt.vals, shape=(nnz,)), o.vals())
    assert np.array_equal(np.ctypeslib.as_array(tt.vals, shape=(nnz,)), val)
    L.splatt_tt_free(p)


def test_layout_switch_rejects_unknown_values():
    L = ps.lib()
    assert L.splatt_amd_set_mttkrp_layout(ps.LAYOUT_CSF) == ps.SUCCESS
    assert L.splatt_amd_set_mttkrp_layout(ps.LAYOUT_AUTO) == ps.SUCCESS
    assert L.splatt_amd_set_mttkrp_layout(7) == ps.ERROR_BADINPUT


def test_coord_load_and_accessors():
    L = ps.lib()
    c = L.splatt_alloc_coord(3, 1)
    assert L.splatt_coord_load(tensor_path("small.tns").encode(), c) == ps.SUCCESS
    L.splatt_coord_get_nnz.restype = C.c_uint64
    L.splatt_coord_get_modes.restype = C.c_uint64
    assert L.splatt_coord_get_nnz(c) == 6 and L.splatt_coord_get_modes(c) == 3
    L.splatt_free_coord(c)


def test_mode_order_matches_oracle():
    L = ps.lib()
    rng = np.random.default_rng(0)
    for trial in range(200):
        n = int(rng.integers(2, 9))
        dims = rng.integers(1, 6, n).astype(np.uint64)
        for typ in range(4):
            mode = int(rng.integers(0, n))
            perm = np.zeros(n, dtype=np.uint64)
            L.splatt_csf_find_mode_order(dims.ctypes.data_as(ps.u64p), n, typ, mode,
                                         perm.ctypes.data_as(ps.u64p))
            assert list(perm) == orc.mode_order(list(dims), typ, mode)


def test_remove_empty_matches_oracle():
    L = ps.lib()
    rng = np.random.default_rng(4)
    dims = [30, 20, 25]
    nnz = 120
    inds = [rng.choice(np.arange(0, d, 2), nnz) for d in dims]
    vals = rng.random(nnz)
    p = L.splatt_alloc_coord(3, nnz)
    for m in range(3):
        p[0].dims[m] = dims[m]
        np.ctypeslib.as_array(p[0].ind[m], shape=(nnz,))[:] = inds[m]
    np.ctypeslib.as_array(p[0].vals, shape=(nnz,))[:] = vals
    removed = L.splatt_tt_remove_empty(p)
    o = orc.Coo.from_arrays(inds, vals, dims)
    assert removed == o.remove_empty()
    for m in range(3):
        assert int(p[0].dims[m]) == o.dims[m]
        assert np.array_equal(np.ctypeslib.as_array(p[0].ind[m], shape=(nnz,)), o.inds()[m])
        # indmap keeps local -> global
        im = np.ctypeslib.as_array(p[0].indmap[m], shape=(o.dims[m],))
        assert np.array_equal(im, np.unique(inds[m]))
    L.splatt_free_coord(p)


def test_mttkrp_ws_map():
    # host-only tensor header: the map and privatisation rule of
    # src/mttkrp.c:1831-1929 need only dims, nnz and dim_iperm
    arr = (ps.SplattCsf * 3)()
    dims = [100, 7, 50]
    for i in range(3):
        arr[i].nnz = 10000
        arr[i].nmodes = 3
        for m in range(3):
            arr[i].dims[m] = dims[m]
    perm = orc.mode_order(dims, 0)       # [1, 2, 0]
    for l in range(3):
        arr[0].dim_perm[l] = perm[l]
        arr[0].dim_iperm[perm[l]] = l
    L = ps.lib()
    for kind, want in ((ps.CSF_ONEMODE, [0, 0, 0]), (ps.CSF_TWOMODE, [1, 0, 0]),
                       (ps.CSF_ALLMODE, [0, 1, 2])):
        o = ps.Opts(csf_alloc=kind, nthreads=8)
        ws = L.splatt_mttkrp_alloc_ws(arr, 16, o.ptr)
        assert [int(ws[0].mode_csf_map[m]) for m in range(3)] == want
        # privatise iff dims*T <= 0.02*nnz: only mode 1 (7*8 = 56 <= 200)
        assert [bool(ws[0].is_privatized[m]) for m in range(3)] == [False, True, False]
        L.splatt_mttkrp_free_ws(ws)


def _integration_program():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"<!-- program: cpd_example.c -->\s*```c\n(.*?)```", text, re.S)
    assert m, "INTEGRATION.md lost its example program"
    return m.group(1)


def build_integration_program(tmpdir):
    """INTEGRATION.md's C program, compiled against include/ and linked
    with libsplatt.so exactly as the document says (C99 and C++)."""
    import shutil
    import subprocess
    src = os.path.join(str(tmpdir), "cpd_example.c")
    with open(src, "w") as f:
        f.write(_integration_program())
    exe = os.path.join(str(tmpdir), "cpd_example")
    pkg = os.path.join(ROOT, "splatt-stream_amd")
    inc = os.path.join(ROOT, "include")
    for cc, std in (("gcc", "-std=c99"), ("g++", "-std=c++14")):
        if shutil.which(cc) is None:
            pytest.skip("needs the host " + cc)
        lang = ["-x", "c++"] if cc == "g++" else []
        subprocess.run([cc, std, "-Wall", "-Werror"] + lang + [src, "-I" + inc, "-L" + pkg,
                        "-lsplatt", "-Wl,-rpath," + pkg, "-o", exe],
                       check=True, capture_output=True, timeout=120)
    return exe


def test_integration_program_compiles_and_links(tmp_path):
    exe = build_integration_program(tmp_path)
    assert os.path.exists(exe)


def test_comm_info_stubs():
    # tests/api_test.c:95-107 of the reference: the non-MPI communicator
    L = ps.lib()
    ci = L.splatt_alloc_comm_info()
    assert ci
    assert ci[0].world_rank == 0 and ci[0].world_npes == 1
    L.splatt_free_comm_info(ci)
    L.splatt_free_comm_info(None)
