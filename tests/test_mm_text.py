"""The Matrix Market body-line rules and value conversion shared by the host
reader and the GPU parser (spai-gmres-ir_amd/csrc/mmtext.hpp), built for the
host from tests/native/mm_text_shim.cpp and compared with glibc's strtod (the
conversion the reference's reader relies on, src/cpu.cpp:105-161): the same
tokens are accepted and rejected with the same code, and accepted values
have the same bits.  Inputs: hand-picked hard cases (halfway ties,
subnormal and overflow boundaries, 750-digit strings, hexadecimal forms
including glibc's treatment of long subnormal significands, inf/nan forms,
partial tokens) and seeded random tokens."""
import ctypes
import ctypes.util
import math
import os
import random
import struct
import subprocess
from fractions import Fraction

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM = os.path.join(HERE, "native", "mm_text_shim.cpp")

# codes of mmtext.hpp's LineCode
OK, BAD_VALUE, NON_FINITE, INTEGER_FIELD = 0, 6, 7, 8


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("mmtext") / "mmshim.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out, SHIM], check=True)
    lib = ctypes.CDLL(out)
    lib.mm_parse_value.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    lib.mm_parse_line.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_ulonglong,
                                  ctypes.c_ulonglong, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_ulonglong),
                                  ctypes.POINTER(ctypes.c_ulonglong),
                                  ctypes.POINTER(ctypes.c_double)]
    return lib


_libc = ctypes.CDLL(ctypes.util.find_library("c"), use_errno=True)
_strtod = _libc.strtod
_strtod.restype = ctypes.c_double
_strtod.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]


def glibc_code(tok, integer=False):
    """The host reader's rule on a value token, by glibc's strtod."""
    buf = ctypes.create_string_buffer(tok.encode())
    end = ctypes.c_char_p()
    ctypes.set_errno(0)
    v = _strtod(buf, ctypes.byref(end))
    err = ctypes.get_errno()
    consumed = ctypes.cast(end, ctypes.c_void_p).value - ctypes.addressof(buf)
    if consumed == 0 or err == 34:
        return BAD_VALUE, None
    if consumed != len(tok) or not math.isfinite(v):
        return NON_FINITE, None
    if integer and any(c not in "+-0123456789" for c in tok):
        return INTEGER_FIELD, None
    return OK, v


def mine(lib, tok, integer=False):
    b = tok.encode()
    v, fast = ctypes.c_double(), ctypes.c_int()
    c = lib.mm_parse_value(b, len(b), int(integer), ctypes.byref(v), ctypes.byref(fast))
    return c, (v.value if c == OK else None)


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def same(lib, tok, integer=False):
    rc, rv = glibc_code(tok, integer)
    mc, mv = mine(lib, tok, integer)
    return rc == mc and (rc != OK or bits(rv) == bits(mv))


HARD = [
    "1e-310", "4.9406564584124654e-324", "2.4703282292062328e-324", "2.4703282292062327e-324",
    "2.2250738585072014e-308", "2.2250738585072011e-308", "2.2250738585072012e-308", "1e-400",
    "0e-999", "0x1p-1074", "0x1.8p-1074", "0x1p-1075", "0x1.0000000000001p-1022",
    "1.7976931348623157e308", "1.7976931348623159e308", "1e309", "-1e-320", "1e-308",
    "0x1.fffffffffffffp-1023", "0x4276ABf5dE6CC20p-1081", "0x1.00000000000008p-1023",
    "0x1.00000000000018p-1023", "0x1.0000000000000cp-1023", "0", "-0", "+0.0", "00000.0000e5",
    "9007199254740993", "9007199254740992", "18014398509481983", "0.1", "3.14159", "inf", "-inf",
    "INF", "Infinity", "-INFINITY", "infin", "nan", "NaN", "nan(123)", "nan(abc_)", "nan(",
    "nan(a b", "nanx", "0x", "0x.", "0xg", "0x.8", "0x1p", "0x1p+", "0X1P3", "0x1.8",
    "0x.0001p16", "0xffffffffffffffffffffp0", "0x1.fffffffffffff8p1023", "1e", "1e+", "1E-",
    "1e5x", ".5", "5.", ".", "+.", "-.e1", "e5", "+", "-", "1.2.3", "1e5.5", "12abc", "0001",
    "1e0001", "123456789012345678901234567890", "0.000000000000000000000000000001",
    "1e99999999999999999999", "1e-99999999999999999", "0e99999999999999", "7.2057594037927933e16",
    "1e22", "1e23", "8.98846567431158e307", "4.4501477170144023e-308", "4.4501477170144024e-308",
    "1.00000000000000011102230246251565404236316680908203125",
    "1.00000000000000011102230246251565404236316680908203124",
    "1.00000000000000011102230246251565404236316680908203126",
    # the exact midpoint between 0 and the smallest subnormal, and just above it
    "2." + "4703282292062327208828439643411068618252990130716238221279284125033775363510437593264991818081799618989828234772285886546332835517796989819938739800539093906315035659515570226392290858392449105184435931802849936536152500319370457678249219365623669863658480757001585769269903706311928279558551332927834338409351978015531246597263579574622766465272827220056374006485499977096599470454020828166226237857393450736339007967761930577506740176324673600968951340535537458516661134223766678604162159680461914467291840300530057530849048765391711386591646239524912623653881879636239373280423891018672348497668235089863388587925628302755995657524455507255189313690836254779186948667994968324049705821028513185451396213837722826145437693412532098591327667236328125e-324",
    "2." + "4703282292062327208828439643411068618252990130716238221279284125033775363510437593264991818081799618989828234772285886546332835517796989819938739800539093906315035659515570226392290858392449105184435931802849936536152500319370457678249219365623669863658480757001585769269903706311928279558551332927834338409351978015531246597263579574622766465272827220056374006485499977096599470454020828166226237857393450736339007967761930577506740176324673600968951340535537458516661134223766678604162159680461914467291840300530057530849048765391711386591646239524912623653881879636239373280423891018672348497668235089863388587925628302755995657524455507255189313690836254779186948667994968324049705821028513185451396213837722826145437693412532098591327667236328125001e-324",
]


def test_hard_cases(shim):
    bad = [t for t in HARD if not same(shim, t)]
    assert not bad, bad[:5]
    for t in ["12", "+12", "-3", "1.5", "1e3", "0x10", "inf"]:
        assert same(shim, t, integer=True), t


def _random_tokens(rng, count):
    for _ in range(count):
        r = rng.random()
        if r < 0.3:
            x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
            if not math.isfinite(x):
                continue
            yield repr(x) if rng.random() < 0.5 else "%.*e" % (rng.randint(0, 25), x)
        elif r < 0.55:
            nd = rng.choice([1, 2, 5, 10, 15, 16, 17, 18, 19, 20, 21, 25, 40])
            d = "".join(rng.choice("0123456789") for _ in range(nd))
            pos = rng.randint(0, nd)
            t = d[:pos] + ("." if rng.random() < 0.7 else "") + d[pos:]
            if rng.random() < 0.8:
                t += "e%d" % rng.randint(-350, 330)
            yield (rng.choice("+-") + t) if rng.random() < 0.3 else t
        elif r < 0.75:  # exact decimal midpoints of adjacent doubles, and their neighbours
            x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(63)))[0]
            y = math.nextafter(x, math.inf)
            if not (math.isfinite(x) and math.isfinite(y)):
                continue
            mid = (Fraction(x) + Fraction(y)) / 2
            k = mid.denominator.bit_length() - 1
            num = mid.numerator * 5 ** k + (rng.choice([-1, 1]) if rng.random() < 0.3 else 0)
            yield "%de-%d" % (num, k)
        elif r < 0.9:  # hexadecimal, including long subnormal significands
            nh = rng.randint(1, 20)
            h = "".join(rng.choice("0123456789abcdefABCDEF") for _ in range(nh))
            pos = rng.randint(0, nh)
            t = "0x" + h[:pos] + ("." if rng.random() < 0.5 else "") + h[pos:]
            yield t + ("p%d" % rng.randint(-1140, 1100) if rng.random() < 0.8 else "")
        else:
            e = rng.choice([-308, -307, -309, -320, -323, -324, -325, 307, 308, 309])
            yield "%d.%de%d" % (rng.randint(1, 9), rng.getrandbits(60), e)


def test_random_tokens_match_glibc(shim):
    rng = random.Random(0x5EED)
    bad = [t for t in _random_tokens(rng, 60000) if not same(shim, t)]
    assert not bad, bad[:5]


def test_entry_lines(shim):
    """Line rules: field count, index checks before the value, precedence."""
    def line(text, rows=10, cols=10, integer=False):
        b = text.encode()
        r, c, v = ctypes.c_ulonglong(), ctypes.c_ulonglong(), ctypes.c_double()
        code = shim.mm_parse_line(b, len(b), rows, cols, int(integer), ctypes.byref(r),
                                  ctypes.byref(c), ctypes.byref(v))
        return code, r.value, c.value, v.value

    assert line("3 4 2.5") == (0, 2, 3, 2.5)
    assert line("  3\t4   -1e1 \r")[0] == 0
    assert line("3 4")[0] == 1                      # Malformed coordinate entry
    assert line("3 4 5 6")[0] == 2                  # Extra field
    assert line("3 x 5")[0] == 3                     # not a digit string
    assert line("-3 4 5")[0] == 3
    assert line("99999999999999999999 4 5")[0] == 4  # beyond size_t
    assert line("0 4 5")[0] == 5                     # one-based
    assert line("11 4 5")[0] == 5
    assert line("3 4 abc")[0] == 6
    assert line("3 4 1e400")[0] == 6
    assert line("3 4 1.5x")[0] == 7
    assert line("3 4 nan")[0] == 7
    assert line("3 4 1.5", integer=True)[0] == 8
    assert line("x 0 abc")[0] == 3                   # the row index is checked first
    assert line("0 x abc")[0] == 3                   # both indices before the range
