"""The server's per-stream generator (CPU only): its factor values equal
the ones glibc's random_r gives from the same 128-byte state (initstate_r
with the stream's seed), two draws per value as src/util.c:15-40 uses them
(magnitude 3 * r / RAND_MAX, negative when the second draw is even)."""
import ctypes as C

import numpy as np
import pytest

import pysplatt as ps

RAND_MAX = 2147483647


class _RandomData(C.Structure):
    _fields_ = [("fptr", C.c_void_p), ("rptr", C.c_void_p), ("state", C.c_void_p),
                ("rand_type", C.c_int), ("rand_deg", C.c_int), ("rand_sep", C.c_int),
                ("end_ptr", C.c_void_p)]


def _glibc_values(seed, n):
    libc = C.CDLL("libc.so.6")
    rd = _RandomData()
    st = C.create_string_buffer(128)
    assert libc.initstate_r(C.c_uint(seed), st, C.c_size_t(128), C.byref(rd)) == 0
    out = np.empty(n)
    a = C.c_int32()
    b = C.c_int32()
    for i in range(n):
        libc.random_r(C.byref(rd), C.byref(a))
        libc.random_r(C.byref(rd), C.byref(b))
        v = 3.0 * (a.value / RAND_MAX)
        out[i] = -v if b.value % 2 == 0 else v
    return out


@pytest.mark.parametrize("seed,n", [(1, 17), (7, 1000), (123456, 40000)])
def test_stream_rand_equals_random_r(seed, n):
    L = ps.lib()
    L.spd_serve_rand_fill.restype = None
    L.spd_serve_rand_fill.argtypes = [C.c_uint, C.c_void_p, C.c_uint64]
    got = np.empty(n)
    L.spd_serve_rand_fill(seed, got.ctypes.data, n)
    assert np.array_equal(got, _glibc_values(seed, n))
