"""Writes pow5_table.hpp: 128-bit normalized powers of five for the
decimal-to-binary conversion of the Matrix Market value parser.

For q in [-342, 308], entry q holds the 128 most significant bits of 5^q,
scaled into [2^127, 2^128): floor for q >= 0 (exact when 5^q has at most 128
bits), and for q < 0 the quotient 2^(b) / 5^-q rounded up, b chosen so that
it lands in [2^127, 2^128).  Exact integer arithmetic (Python ints).

Usage: python gen_pow5.py > pow5_table.hpp
"""

QMIN, QMAX = -342, 308


def entry(q: int) -> int:
    if q >= 0:
        v = 5 ** q
        bits = v.bit_length()
        return v << (128 - bits) if bits <= 128 else v >> (bits - 128)
    d = 5 ** (-q)
    b = d.bit_length() + 127  # 2^b / d in [2^127, 2^128)
    t = -(-(1 << b) // d)     # ceiling
    if t >= 1 << 128:         # (only if d were a power of two)
        t >>= 1
    return t


def floor_log2_pow5(q: int) -> int:
    if q >= 0:
        return (5 ** q).bit_length() - 1
    return -((5 ** (-q)).bit_length())  # 5^-|q| lies in (2^-bits, 2^-(bits-1))


def main() -> None:
    print("// Generated by gen_pow5.py: 128-bit normalized powers of five, q = %d .. %d." % (QMIN, QMAX))
    print("// Row q - kPow5Min: {high 64 bits, low 64 bits}.")
    print("#pragma once")
    print("#include <cstdint>")
    print("namespace bspai_mm {")
    print("constexpr int kPow5Min = %d, kPow5Max = %d;" % (QMIN, QMAX))
    print("constexpr std::uint64_t kPow5Table[%d][2] = {" % (QMAX - QMIN + 1))
    for q in range(QMIN, QMAX + 1):
        v = entry(q)
        assert (1 << 127) <= v < (1 << 128), q
        print("    {0x%016xULL, 0x%016xULL}," % (v >> 64, v & ((1 << 64) - 1)))
    print("};")
    print("// floor(log2(5^q)) for the same rows.")
    print("constexpr short kPow5Log2[%d] = {" % (QMAX - QMIN + 1))
    vals = []
    for q in range(QMIN, QMAX + 1):
        vals.append(str(floor_log2_pow5(q)))
    for i in range(0, len(vals), 16):
        print("    " + ", ".join(vals[i:i + 16]) + ",")
    print("};")
    print("} // namespace bspai_mm")


if __name__ == "__main__":
    main()
