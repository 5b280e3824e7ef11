"""Pins the oracle's IO, mode ordering and CSF construction.

What pins what:
  or_tns_read      the reference fixtures (tests/tensors/*.tns): hand-read
                   small tensors, 0/1-index detection, dims of med*.tns as
                   the reference's csf_test reads them; numpy.loadtxt parse.
  or_remove_empty  numpy.unique relabelling on tensors with planted gaps.
  or_mode_order    the reference's csf_test cases (tests/csf_test.c:63-162)
                   and hand-worked tie cases.
  or_csf_build     a hand-derived CSF of small.tns (tests/golden/
                   small_csf.json); an independent nested-groupby trie on
                   random tensors; the reference's csf_test invariants
                   (vals/leaf ids equal the sorted COO, fids[0] NULL untiled,
                   nfibs[0] = root dim, frobsq = sum v^2); DENSETILE tile
                   contents against a direct bucketing.
"""
import itertools
import json
import os

import numpy as np
import pytest

from conftest import DATASETS, GOLDEN, tensor_path
import oracle_py as orc


def test_read_small():
    c = orc.Coo.read(tensor_path("small.tns"))
    assert c.nmodes == 3 and c.nnz == 6 and c.dims == [2, 3, 2]
    ind = c.inds()
    got = sorted(zip(ind[0], ind[1], ind[2], c.vals()))
    want = sorted([(1, 1, 1, 6.0), (0, 2, 0, 2.0), (0, 1, 0, 1.0), (1, 2, 0, 5.0),
                   (0, 0, 1, 3.0), (1, 0, 0, 4.0)])
    assert got == want


def test_read_zero_index_detection():
    one = orc.Coo.read(tensor_path("small4.tns"))
    zero = orc.Coo.read(tensor_path("small4_zeroidx.tns"))
    # the two files hold the same tensor; the last mode is 0-indexed in one
    assert one.dims == zero.dims == [2, 2, 3, 3]
    for a, b in zip(one.inds(), zero.inds()):
        assert np.array_equal(a, b)
    assert np.array_equal(one.vals(), zero.vals())


@pytest.mark.parametrize("name", DATASETS)
def test_read_matches_loadtxt(name):
    c = orc.Coo.read(tensor_path(name))
    raw = np.loadtxt(tensor_path(name), comments="#", ndmin=2)
    n = raw.shape[1] - 1
    assert c.nmodes == n and c.nnz == raw.shape[0]
    for m in range(n):
        col = raw[:, m].astype(np.uint64)
        off = 0 if col.min() == 0 else 1
        assert np.array_equal(c.inds()[m], col - off)
        assert c.dims[m] == int(col.max()) + (1 - off)
    assert np.array_equal(c.vals(), raw[:, n])


def test_med_dims_known():
    assert orc.Coo.read(tensor_path("med.tns")).dims == [2425, 10816, 29567]
    assert orc.Coo.read(tensor_path("med4.tns")).dims == [60114, 83826, 23034, 1132]


def test_remove_empty():
    rng = np.random.default_rng(3)
    dims = [40, 30, 20]
    nnz = 200
    # draw from a subset of each mode so some slices are empty
    keep = [np.sort(rng.choice(d, d // 2, replace=False)) for d in dims]
    inds = [rng.choice(k, nnz) for k in keep]
    vals = rng.random(nnz)
    c = orc.Coo.from_arrays(inds, vals, dims)
    removed = c.remove_empty()
    new = c.inds()
    total = 0
    for m in range(3):
        u, inv = np.unique(inds[m], return_inverse=True)
        assert np.array_equal(new[m], inv.astype(np.uint64))
        assert c.dims[m] == len(u)
        total += dims[m] - len(u)
    assert removed == total


def test_mode_order_reference_cases():
    # tests/csf_test.c:93-110: minus-one puts the mode first, the rest by size
    dims = [10, 9, 8, 7, 0]
    for m in range(5):
        p = orc.mode_order(dims, 3, m)
        assert p[0] == m
        rest = [dims[x] for x in p[1:]]
        assert rest == sorted(rest)
    # tests/csf_test.c:113-132: in-order minus-one
    for m in range(5):
        p = orc.mode_order(dims, 2, m)
        assert p[0] == m and p[1:] == sorted(p[1:])
    # ties: the lower mode index comes first, for either direction
    assert orc.mode_order([5, 5, 3], 0) == [2, 0, 1]
    assert orc.mode_order([3, 5, 5], 1) == [1, 2, 0]
    assert orc.mode_order([4, 4, 4, 4], 0) == [0, 1, 2, 3]
    assert orc.mode_order([4, 4, 4, 4], 1) == [0, 1, 2, 3]
    assert orc.mode_order([2, 3, 2], 3, 1) == [1, 0, 2]


def _check_against(arrays, g):
    t = arrays["tiles"][0]
    assert arrays["perm"] == g["perm"]
    assert t["nfibs"] == g["nfibs"]
    for l in range(3):
        if l < 2:
            assert list(t["fptr"][l]) == g["fptr"][l]
        if g["fids"][l] is None:
            assert t["fids"][l] is None
        else:
            assert list(t["fids"][l]) == g["fids"][l]
    assert list(t["vals"]) == g["vals"]


def test_small_hand_derived():
    g = json.load(open(os.path.join(GOLDEN, "small_csf.json")))
    c = orc.Coo.read(tensor_path("small.tns"))
    for key, (typ, mode) in (("onemode", (0, 0)), ("minusone_1", (3, 1)),
                             ("minusone_2", (3, 2))):
        perm = orc.mode_order(c.dims, typ, mode)
        _check_against(orc.Csf(c, perm).arrays(), g[key])


def _trie(inds, vals, perm):
    """CSF by nested grouping of the lexicographically sorted nonzeros."""
    N = len(perm)
    rows = sorted(zip(*[inds[p] for p in perm], range(len(vals))))
    nfibs = [0] * N
    fptr = [[] for _ in range(N - 1)]
    fids = [[] for _ in range(N)]
    svals = []

    # leaf level: every nonzero is its own node even when coordinates repeat
    def walk_leaf(group, level):
        if level == N - 1:
            for r in group:
                fids[level].append(int(r[level]))
                nfibs[level] += 1
                svals.append(vals[r[-1]])
            return
        for key, sub in itertools.groupby(group, key=lambda r: r[level]):
            sub = list(sub)
            fids[level].append(int(key))
            fptr[level].append(nfibs[level + 1])
            nfibs[level] += 1
            walk_leaf(sub, level + 1)

    walk_leaf(rows, 0)
    for l in range(N - 1):
        fptr[l].append(nfibs[l + 1])
    return nfibs, fptr, fids, svals


@pytest.mark.parametrize("seed,dims,nnz", [(1, [7, 5, 6], 60), (2, [4, 9, 3, 5], 150),
                                           (3, [3, 4, 2, 3, 5], 200), (4, [50, 2, 40], 300)])
def test_build_matches_trie(seed, dims, nnz):
    rng = np.random.default_rng(seed)
    cells = rng.choice(int(np.prod(dims)), nnz, replace=False)
    inds = list(np.unravel_index(cells, dims))
    vals = rng.random(nnz)
    c = orc.Coo.from_arrays(inds, vals, dims)
    for typ in (0, 1, 2, 3):
        for mode in range(len(dims)):
            perm = orc.mode_order(dims, typ, mode)
            a = orc.Csf(c, perm).arrays()["tiles"][0]
            nf, fp, fi, sv = _trie(inds, vals, perm)
            assert a["nfibs"] == nf
            for l in range(len(dims) - 1):
                assert list(a["fptr"][l]) == fp[l]
            assert a["fids"][0] is None
            for l in range(1, len(dims)):
                assert list(a["fids"][l]) == fi[l]
            assert np.array_equal(a["vals"], np.array(sv))


@pytest.mark.parametrize("name", DATASETS)
def test_reference_invariants(name):
    c = orc.Coo.read(tensor_path(name))
    perm = orc.mode_order(c.dims, 0)
    csf = orc.Csf(c, perm, reorder=True)
    a = csf.arrays()
    t = a["tiles"][0]
    N = c.nmodes
    # tests/csf_test.c:31-61: one tile; the COO left sorted equals the leaves
    assert a["ntiles"] == 1 and t["fids"][0] is None
    assert np.array_equal(t["fids"][N - 1], c.inds()[perm[N - 1]])
    assert np.array_equal(t["vals"], c.vals())
    # lexicographic order of the reordered COO
    keys = np.lexsort([c.inds()[perm[l]] for l in reversed(range(N))])
    assert np.array_equal(keys, np.arange(c.nnz))
    # tests/csf_test.c:166-203
    assert abs(csf.frobsq() - float(np.sum(c.vals() ** 2))) <= 1e-9 * csf.frobsq()


def test_nfibs_root_is_dim_med4():
    c = orc.Coo.read(tensor_path("med4.tns"))
    perm = orc.mode_order(c.dims, 0)
    a = orc.Csf(c, perm).arrays()
    assert a["tiles"][0]["nfibs"][0] == c.dims[perm[0]]


@pytest.mark.parametrize("tilelevel,nthreads", [(1, 3), (2, 2), (3, 2), (1, 1)])
def test_densetile_matches_bucketing(tilelevel, nthreads):
    rng = np.random.default_rng(9)
    dims = [9, 11, 7]
    nnz = 250
    cells = rng.choice(int(np.prod(dims)), nnz, replace=False)
    inds = list(np.unravel_index(cells, dims))
    vals = rng.random(nnz)
    c = orc.Coo.from_arrays(inds, vals, dims)
    perm = orc.mode_order(dims, 0)
    a = orc.Csf(c, perm, tile=1, tilelevel=tilelevel, nthreads=nthreads).arrays()
    N = 3
    depth = {perm[l]: l for l in range(N)}
    tdims = [nthreads if depth[m] >= N - min(tilelevel, N) else 1 for m in range(N)]
    assert a["tile_dims"] == tdims
    ntiles = int(np.prod(tdims))
    assert a["ntiles"] == ntiles
    tsize = [max(dims[m] // tdims[m], 1) for m in range(N)]
    tid = np.zeros(nnz, dtype=np.int64)
    for m in range(N):
        tid = tid * tdims[m] + np.minimum(inds[m] // tsize[m], tdims[m] - 1)
    for t in range(ntiles):
        sel = tid == t
        tile = a["tiles"][t]
        if not sel.any():
            assert tile["vals"] is None and list(tile["fptr"][0]) == [0, 0]
            continue
        sub = [x[sel] for x in inds]
        nf, fp, fi, sv = _trie(sub, vals[sel], perm)
        assert tile["nfibs"] == nf
        for l in range(N - 1):
            assert list(tile["fptr"][l]) == fp[l]
        for l in range(N):
            if l == 0 and ntiles == 1:
                assert tile["fids"][0] is None
            else:
                assert list(tile["fids"][l]) == fi[l]
        assert np.array_equal(tile["vals"], np.array(sv))
