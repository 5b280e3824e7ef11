"""The text reader's field conversions (the code the GPU runs for every line
of a .tns file), run on the host and checked against the C library's
strtod/strtoull, which the reference's reader calls (src/io.c:273-348).

A conversion may decline a field (the reader then converts that line with
the C library); when it accepts one, its result must be the C library's bit
for bit. The cases are chosen so that a plausible mistake fails one of them:
a wrong table entry or exponent (random exponents over the whole range),
rounding of exact halfway cases (the 2^53 + 1 family), the carry into the
next binade, the ends of the normal range, signs and signed zero, and
fields that strtod reads further than the device accepts (hexadecimal,
infinities)."""
import ctypes as C
import math
import struct

import numpy as np
import pytest

import pysplatt as ps

_libc = C.CDLL(None)
_libc.strtod.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]
_libc.strtod.restype = C.c_double
_libc.strtoull.argtypes = [C.c_char_p, C.POINTER(C.c_char_p), C.c_int]
_libc.strtoull.restype = C.c_ulonglong


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _convert(fields, kind):
    """spd_decimal_convert over a list of str fields: (values, ok)."""
    enc = [f.encode() for f in fields]
    buf = b"".join(enc)
    off = np.zeros(len(enc) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(e) for e in enc])
    n = len(enc)
    vals = np.zeros(n, dtype=np.float64)
    idx = np.zeros(n, dtype=np.uint64)
    ok = np.zeros(n, dtype=np.int32)
    rc = ps.lib().spd_decimal_convert(buf, off.ctypes.data_as(ps.u64p), n, kind,
                                      vals.ctypes.data_as(ps.f64p),
                                      idx.ctypes.data_as(ps.u64p),
                                      ok.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == 0
    return (vals if kind == 0 else idx), ok.astype(bool)


def _strtod(f):
    return _libc.strtod(f.encode(), None)


def _check_values(fields, must_accept=True):
    got, ok = _convert(fields, 0)
    bad = []
    for f, g, k in zip(fields, got, ok):
        if not k:
            if must_accept:
                bad.append((f, "declined"))
            continue
        want = _strtod(f)
        if _bits(float(g)) != _bits(want):
            bad.append((f, float(g), want))
    assert bad == [], bad[:10]
    return ok


def test_shortest_round_trip_strings():
    rng = np.random.default_rng(11)
    # doubles spread over the whole normal range, both signs
    b = rng.integers(0x0010000000000000, 0x7FEFFFFFFFFFFFFF, size=40000, dtype=np.uint64)
    x = b.view(np.float64) * np.where(rng.random(b.size) < 0.5, -1.0, 1.0)
    fields = [repr(float(v)) for v in x]
    _check_values(fields)
    # what each field means must also be what Python reads
    got, _ = _convert(fields, 0)
    assert np.array_equal(got, x)


@pytest.mark.parametrize("fmt", ["%.17g", "%.16e", "%.6e", "%.3f", "%g", "%.12f", "%.19e"])
def test_printf_shapes(fmt):
    rng = np.random.default_rng(sum(map(ord, fmt)))
    x = np.concatenate([rng.standard_normal(8000), rng.lognormal(0, 20, 8000),
                        rng.uniform(-1e6, 1e6, 8000)])
    fields = [fmt % v for v in x]
    # fixed-point forms of large values carry more than 19 significant
    # digits: those may be declined
    ok = _check_values(fields, must_accept=False)
    if fmt in ("%.17g", "%.16e", "%.6e", "%g"):
        assert ok.all()
    elif fmt == "%.19e":
        assert not ok.any()          # 20 significant digits
    else:
        assert ok.mean() > 0.8


def test_random_decimals_every_exponent():
    # significands of 1..19 digits at every decimal exponent the table
    # covers, written with the point in random places
    rng = np.random.default_rng(5)
    fields = []
    for q in range(-342, 309):
        for _ in range(24):
            nd = int(rng.integers(1, 20))
            s = str(int(rng.integers(1, 10))) + "".join(map(str, rng.integers(0, 10, nd - 1)))
            pt = int(rng.integers(0, nd + 1))
            mant = s[:pt] + "." + s[pt:] if pt < nd else s
            e = q + (nd - pt if pt < nd else 0)
            fields.append("%s%se%d" % ("-" if rng.random() < 0.3 else "", mant, e))
    got, ok = _convert(fields, 0)
    nacc = 0
    declined_normal = []
    for f, g, k in zip(fields, got, ok):
        want = _strtod(f)
        if k:
            nacc += 1
            assert _bits(float(g)) == _bits(want), f
        elif want != 0.0 and math.isfinite(want) and abs(want) >= 2.2250738585072014e-308:
            declined_normal.append(f)
    # only results outside the normal range are left to the C library (an
    # undecided product would be too, but has odds of about 2^-64)
    assert declined_normal == []
    assert nacc > 0.9 * len(fields)


def test_halfway_and_carry_cases():
    cases = [
        # exactly halfway between two doubles: round to even
        "9007199254740993", "9007199254740995", "9007199254740997",
        "18014398509481989", "18014398509481990", "18014398509481991",
        "4503599627370496.5", "4503599627370497.5",
        "1.0000000000000001110", "1.000000000000000111",
        # 2^64 - 1 and neighbours: the significand fills 64 bits
        "18446744073709551615", "18446744073709551614", "9999999999999999999",
        # carries into the next binade
        "9007199254740991", "9007199254740992", "1.7976931348623157e308",
        "0.99999999999999999", "1.9999999999999999", "3.9999999999999998",
        # ends of the normal range
        "2.2250738585072014e-308", "2.2250738585072012e-308", "2.2250738585072015e-308",
        "1.7976931348623157e+308", "8.98846567431158e307",
        # classic hard cases
        "1e23", "8.41e21", "7.038531e-26", "9.7656250e-4", "0.1", "0.3",
        "3.14159265358979323", "2.718281828459045", "1e-300", "1e300", "1e-307",
        "123456789012345678e-20", "1234567890123456789e280", "1e22", "1e-22",
        # signs and zeros
        "-0", "+0", "-0.0", "0e0", "0.000000", "-0e-999", "+1.5", "-.5", "5.",
    ]
    cases = [c for c in cases if c]
    ok = _check_values(cases, must_accept=False)
    accepted = {c for c, k in zip(cases, ok) if k}
    for c in ("9007199254740993", "9007199254740995", "1e23", "0.1", "-0", "-.5", "5.",
              "2.2250738585072014e-308", "1.7976931348623157e308", "9999999999999999999",
              "0.99999999999999999", "1.000000000000000111"):
        assert c in accepted, c
    got, _ = _convert(["-0"], 0)
    assert math.copysign(1.0, got[0]) == -1.0
    got, _ = _convert(["9007199254740993"], 0)
    assert got[0] == 9007199254740992.0


def test_declined_fields():
    # fields strtod reads differently or that leave the normal range go to
    # the C library; none may be accepted with another meaning
    fields = ["0x1p3", "-0x10", "inf", "-Infinity", "nan", "NAN(123)", "1e", "1e+",
              "1.5abc", "", ".", "-", "+", "e5", "1e400", "-1e400", "4.9e-324",
              "2.2250738585072011e-308", "1e-320", "12345678901234567890",
              "1.2345678901234567890", "1,5", "1_000", "\x001"]
    got, ok = _convert(fields, 0)
    assert not ok.any(), [f for f, k in zip(fields, ok) if k]


def test_blanks_around_values():
    fields = ["  1.5", "\t-2", " 3e2 ", "4\r", "5\t\t", "\v6"]
    got, ok = _convert(fields, 0)
    assert ok.all()
    assert list(got) == [1.5, -2.0, 300.0, 4.0, 5.0, 6.0]


def test_index_fields():
    rng = np.random.default_rng(2)
    vals = list(rng.integers(0, 2 ** 63, size=5000, dtype=np.uint64)) + [0, 1, 10 ** 19 - 1]
    fields = [str(int(v)) for v in vals]
    fields += ["007", "+12", "  42", "\t9 ", "1\r"]
    got, ok = _convert(fields, 1)
    assert ok.all()
    for f, g in zip(fields, got):
        assert int(g) == _libc.strtoull(f.encode(), None, 10), f
    # strtoull reads these differently (sign, 20 digits, other characters)
    bad = ["-1", "12345678901234567890", "12a", "1.0", "", "+", "0x10", "1e3"]
    got, ok = _convert(bad, 1)
    assert not ok.any(), [f for f, k in zip(bad, ok) if k]
