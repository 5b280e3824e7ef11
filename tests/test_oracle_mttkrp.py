"""Pins the oracle's MTTKRP.

  or_mttkrp_coo    against the textbook definition on a densified tensor
                   (numpy einsum of the dense tensor with the factors).
  or_mttkrp_csf    against or_mttkrp_coo on the reference fixtures and on
                   random tensors, for every output level (root, internal,
                   leaf), for tiled and untiled tensors and 1 or 4 threads:
                   the reference's own known-answer procedure
                   (tests/mttkrp_test.c:36-88, tolerance 1e-10 in FP64).
"""
import numpy as np
import pytest

from conftest import DATASETS, tensor_path
import oracle_py as orc

LETTERS = "abcdefgh"


def _dense_mttkrp(inds, vals, dims, mats, mode):
    X = np.zeros(dims)
    np.add.at(X, tuple(inds), vals)
    N = len(dims)
    spec = LETTERS[:N] + "".join("," + LETTERS[m] + "r" for m in range(N) if m != mode)
    spec += "->" + LETTERS[mode] + "r"
    return np.einsum(spec, X, *[mats[m] for m in range(N) if m != mode])


@pytest.mark.parametrize("dims,nnz", [([5, 4, 6], 40), ([3, 4, 5, 2], 70),
                                      ([2, 3, 2, 3, 2], 50)])
def test_coo_matches_dense(dims, nnz):
    rng = np.random.default_rng(len(dims))
    cells = rng.choice(int(np.prod(dims)), nnz, replace=False)
    inds = list(np.unravel_index(cells, dims))
    vals = rng.standard_normal(nnz)
    mats = [rng.standard_normal((d, 3)) for d in dims]
    c = orc.Coo.from_arrays(inds, vals, dims)
    for mode in range(len(dims)):
        want = _dense_mttkrp(inds, vals, dims, mats, mode)
        assert np.allclose(orc.mttkrp_coo(c, mats, mode), want, rtol=1e-13, atol=1e-13)
        got32 = orc.mttkrp_coo(c, [m.astype(np.float32) for m in mats], mode, f32=True)
        assert np.allclose(got32, want, rtol=1e-4, atol=1e-4)


def _relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", DATASETS)
@pytest.mark.parametrize("nthreads", [1, 4])
def test_csf_matches_coo_fixtures(name, nthreads):
    c = orc.Coo.read(tensor_path(name))
    rng = np.random.default_rng(5)
    mats = [rng.random((d, 3)) for d in c.dims]
    N = c.nmodes
    for mode in range(N):
        gold = orc.mttkrp_coo(c, mats, mode)
        # one tensor ordered small-first (root/internal/leaf outputs) and the
        # all-mode tensor rooted at `mode`
        for perm in (orc.mode_order(c.dims, 0), orc.mode_order(c.dims, 3, mode)):
            got = orc.Csf(c, perm).mttkrp(mats, mode, nthreads=nthreads)
            assert np.max(np.abs(got - gold)) <= 1e-10


@pytest.mark.parametrize("tilelevel", [1, 2, 3])
def test_csf_tiled_matches_coo(tilelevel):
    c = orc.Coo.read(tensor_path("med.tns"))
    rng = np.random.default_rng(6)
    mats = [rng.random((d, 4)) for d in c.dims]
    perm = orc.mode_order(c.dims, 0)
    t = orc.Csf(c, perm, tile=1, tilelevel=tilelevel, nthreads=3)
    for mode in range(3):
        assert np.max(np.abs(t.mttkrp(mats, mode, nthreads=3)
                             - orc.mttkrp_coo(c, mats, mode))) <= 1e-10


def test_csf_f32():
    c = orc.Coo.read(tensor_path("med4.tns"))
    rng = np.random.default_rng(8)
    mats = [rng.random((d, 8)).astype(np.float32) for d in c.dims]
    for mode in range(4):
        gold = orc.mttkrp_coo(c, mats, mode, f32=True)
        got = orc.Csf(c, orc.mode_order(c.dims, 0)).mttkrp(mats, mode, nthreads=2, f32=True)
        assert _relerr(got.astype(np.float64), gold.astype(np.float64)) <= 1e-5
