"""Shared test setup: markers, paths, and loading the built libraries.

GPU tests are marked `gpu`; everything else runs on the CPU. Libraries are
built in-tree by __graft_entry__.build(); if they are missing here they are
built once (CPU-side compile, no GPU needed).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "splatt-stream_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
TENSORS = os.path.join(GOLDEN, "tensors")

for p in (ROOT, PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)

# datasets of the reference's own tests (tests/splatt_test.h:11-17)
DATASETS = ["small.tns", "med.tns", "small4.tns", "med4.tns", "med5.tns"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")


def _ensure_built():
    need = [os.path.join(PKG, "libsplatt.so"), os.path.join(ORACLE, "liboracle.so"),
            os.path.join(PKG, "gen", "libsplatt_gen.so")]
    if not all(os.path.exists(p) for p in need):
        import __graft_entry__
        __graft_entry__.build()


_ensure_built()


def tensor_path(name):
    return os.path.join(TENSORS, name)


@pytest.fixture(scope="session")
def gpu_available():
    import pysplatt
    import ctypes
    L = pysplatt.lib()
    p = L.splatt_amd_dev_alloc(256)
    if not p:
        pytest.fail("GPU tests need a usable GPU: " + pysplatt.last_error())
    L.splatt_amd_dev_free(ctypes.c_void_p(p))
    return True
