// Matrix Market text rules of the body lines and the conversion of a value
// token, shared by the host reader (mmio.cpp) and the GPU parser (mmgpu.hip),
// so that both accept, reject and convert exactly alike.
//
// The rules are the reference reader's (src/cpu.cpp:105-161): a line is data
// unless it is empty after its leading " \t\r" or its first other character
// is '%'; fields are separated by C-locale white space; a data line holds
// exactly three fields; indices are decimal digit strings within size_t and
// one-based within the dimensions; a value token is what strtod accepts, in
// full, finite and without a range error; an integer field allows only the
// characters "+-0123456789".
//
// strtod (glibc, C locale) is restated here, not called: the GPU has none.
//   * grammar: optional sign; "inf"/"infinity"/"nan"/"nan(chars)" in any
//     case; hexadecimal "0x" digits [. digits] [p exponent]; decimal digits
//     [. digits] [e exponent]; the longest such prefix is consumed ("0x"
//     without a hex digit consumes the "0"; an exponent marker without
//     digits is not consumed);
//   * value: correctly rounded, ties to even;
//   * range error: the result overflows, or the exact value is nonzero,
//     below 2^-1022 and not representable (tininess before rounding, as
//     glibc reports it: 2.2250738585072012e-308 rounds up to 2^-1022 and
//     still reports it).
// Decimal conversion: Clinger's exact fast path (at most 2^53 and |q| <= 22),
// then the Eisel-Lemire method (D. Lemire, "Number parsing at a gigabyte per
// second", Software: Practice and Experience 51(8), 2021) on 128-bit powers
// of five from pow5_table.hpp for up to 19 significant digits and a normal
// result, and otherwise an exact big-integer conversion (mm_exact_decimal:
// binary exponent from bit lengths, the 53-bit quotient by restoring
// division, the remainder for rounding and exactness).
#pragma once

#include <cstdint>

#include "pow5_table.hpp"

#ifdef __HIPCC__
#define BSPAI_HD __host__ __device__
#else
#define BSPAI_HD
#endif

namespace bspai_mm {

// Outcome of a body line, in the reference reader's precedence within a line.
enum LineCode : int {
    kLineOk = 0,
    kMalformedEntry = 1,  // fewer than three fields
    kExtraField = 2,      // more than three
    kBadIndex = 3,        // an index that is not a digit string
    kIndexOverflow = 4,   // an index beyond size_t
    kIndexRange = 5,      // an index outside 1 .. dimension
    kBadValue = 6,        // nothing converted, or a range error
    kNonFinite = 7,       // not the whole token, or not finite
    kIntegerField = 8,    // a character outside "+-0123456789" in an integer field
    kNeedsExact = 9,      // (GPU internal) the value needs the exact conversion
};

// The messages of the reference reader for the codes above.
inline const char *line_message(int code) {
    switch (code) {
    case kMalformedEntry:
        return "Malformed coordinate entry";
    case kExtraField:
        return "Extra field in Matrix Market line";
    case kBadIndex:
        return "Expected a nonnegative integer dimension/index";
    case kIndexOverflow:
        return "Dimension/index exceeds size_t";
    case kIndexRange:
        return "Matrix Market index outside one-based dimensions";
    case kBadValue:
        return "Invalid coordinate value";
    case kNonFinite:
        return "Coordinate value must be finite and numeric";
    case kIntegerField:
        return "Integer field requires integer value tokens";
    default:
        return "";
    }
}

BSPAI_HD inline bool field_space(char c) {
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}
BSPAI_HD inline bool blank_lead(char c) { return c == ' ' || c == '\t' || c == '\r'; }

BSPAI_HD inline bool is_data_line(const char *b, const char *e) {
    while (b < e && blank_lead(*b))
        ++b;
    return b < e && *b != '%';
}

// Up to `want` fields of [b, e) into fb/fe; returns the count found, capped at
// want + 1 (one more means "extra field").
BSPAI_HD inline int split_fields(const char *b, const char *e, const char **fb, const char **fe,
                                 int want) {
    int n = 0;
    const char *p = b;
    while (n <= want) {
        while (p < e && field_space(*p))
            ++p;
        if (p == e)
            break;
        const char *s = p;
        while (p < e && !field_space(*p))
            ++p;
        if (n < want) {
            fb[n] = s;
            fe[n] = p;
        }
        ++n;
    }
    return n;
}

// A decimal digit string within 64 bits (size_t): kLineOk, kBadIndex or
// kIndexOverflow (every character is checked before the magnitude).
BSPAI_HD inline int parse_index(const char *b, const char *e, std::uint64_t &v) {
    for (const char *p = b; p < e; ++p)
        if (*p < '0' || *p > '9')
            return kBadIndex;
    std::uint64_t x = 0;
    const std::uint64_t top = ~0ull;
    for (const char *p = b; p < e; ++p) {
        const std::uint64_t d = (std::uint64_t)(*p - '0');
        if (x > (top - d) / 10)
            return kIndexOverflow;
        x = x * 10 + d;
    }
    v = x;
    return kLineOk;
}

// ------------------------------------------------------------ bits helpers
BSPAI_HD inline int clz64(std::uint64_t x) {
    int n = 0;
    if (!(x >> 32)) {
        n += 32;
        x <<= 32;
    }
    if (!(x >> 48)) {
        n += 16;
        x <<= 16;
    }
    if (!(x >> 56)) {
        n += 8;
        x <<= 8;
    }
    if (!(x >> 60)) {
        n += 4;
        x <<= 4;
    }
    if (!(x >> 62)) {
        n += 2;
        x <<= 2;
    }
    if (!(x >> 63))
        n += 1;
    return n;
}

// 64 x 64 -> 128 bits from 32-bit halves.
BSPAI_HD inline void mul64(std::uint64_t a, std::uint64_t b, std::uint64_t &hi, std::uint64_t &lo) {
    const std::uint64_t a0 = a & 0xffffffffu, a1 = a >> 32, b0 = b & 0xffffffffu, b1 = b >> 32;
    const std::uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    const std::uint64_t mid = (p00 >> 32) + (p01 & 0xffffffffu) + (p10 & 0xffffffffu);
    lo = (mid << 32) | (p00 & 0xffffffffu);
    hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

BSPAI_HD inline double make_double(bool neg, std::uint64_t biased, std::uint64_t frac) {
    const std::uint64_t bits = ((std::uint64_t)neg << 63) | (biased << 52) | frac;
    double d;
    __builtin_memcpy(&d, &bits, 8);
    return d;
}
BSPAI_HD inline double inf_of(bool neg) { return make_double(neg, 2047, 0); }
BSPAI_HD inline double nan_of(bool neg) { return make_double(neg, 2047, 1ull << 51); }

// Rounds value = mant * 2^e2lo (mant != 0; sticky: more nonzero bits below
// mant's last) to a double: returns the value and sets range (overflow, or
// tiny and inexact).
BSPAI_HD inline double round_binary(bool neg, std::uint64_t mant, int e2lo, bool sticky,
                                    bool &range) {
    const int len = 64 - clz64(mant);
    const int e2 = e2lo + len - 1; // floor(log2 value) (exact: sticky only adds below)
    int u = e2 - 52;               // exponent of the result's last bit
    if (u < -1074)
        u = -1074;
    std::uint64_t m;
    bool half = false, below = sticky; // bits of value below m's last bit: exactly 1/2? others?
    const int s = u - e2lo;            // bits of mant to drop
    if (s <= 0) {
        m = mant << (-s); // (len + (-s) <= 53)
    } else if (s >= 64) {
        m = 0;
        half = s == 64 ? (mant >> 63) != 0 : false;
        below = below || (s == 64 ? (mant << 1) != 0 : mant != 0);
    } else {
        m = mant >> s;
        const std::uint64_t rest = mant & ((1ull << s) - 1);
        const std::uint64_t h = 1ull << (s - 1);
        half = (rest & h) != 0;
        below = below || (rest & (h - 1)) != 0;
    }
    const bool inexact = half || below;
    if (half && (below || (m & 1)))
        ++m;
    if (m == (1ull << 53)) { // carry into a new binade
        m >>= 1;
        ++u;
    }
    const bool tiny = e2 < -1022;
    if (u + 52 > 1023) {
        range = true;
        return inf_of(neg);
    }
    range = tiny && inexact;
    if (m >= (1ull << 52)) // normal
        return make_double(neg, (std::uint64_t)(u + 52 + 1023), m - (1ull << 52));
    return make_double(neg, 0, m); // subnormal (u == -1074) or zero
}

// ------------------------------------------------------------ big integers
// Little-endian 32-bit limbs, fixed capacity (the exact conversion only).
constexpr int kBigLimbs = 136;
struct Big {
    int n = 0; // limbs in use (no leading zero limb)
    std::uint32_t d[kBigLimbs];
};
BSPAI_HD inline void big_set(Big &a, std::uint64_t v) {
    a.n = 0;
    while (v) {
        a.d[a.n++] = (std::uint32_t)v;
        v >>= 32;
    }
}
BSPAI_HD inline void big_mul_small(Big &a, std::uint32_t m, std::uint32_t add = 0) {
    std::uint64_t carry = add;
    for (int i = 0; i < a.n; ++i) {
        const std::uint64_t t = (std::uint64_t)a.d[i] * m + carry;
        a.d[i] = (std::uint32_t)t;
        carry = t >> 32;
    }
    if (carry)
        a.d[a.n++] = (std::uint32_t)carry;
}
BSPAI_HD inline int big_bits(const Big &a) {
    if (a.n == 0)
        return 0;
    return 32 * (a.n - 1) + (64 - clz64(a.d[a.n - 1]));
}
BSPAI_HD inline void big_shl(Big &a, int s) {
    if (a.n == 0 || s == 0)
        return;
    const int w = s / 32, b = s % 32;
    int n = a.n + w + 1;
    for (int i = n - 1; i >= 0; --i) {
        const int j = i - w;
        std::uint64_t v = 0;
        if (j >= 0 && j < a.n)
            v = (std::uint64_t)a.d[j] << b;
        if (b && j - 1 >= 0 && j - 1 < a.n)
            v |= (std::uint64_t)a.d[j - 1] >> (32 - b);
        a.d[i] = (std::uint32_t)v;
    }
    while (n > 0 && a.d[n - 1] == 0)
        --n;
    a.n = n;
}
BSPAI_HD inline void big_shr(Big &dst, const Big &a, int s) { // dst = a >> s
    const int w = s / 32, b = s % 32;
    dst.n = 0;
    for (int i = 0; i + w < a.n; ++i) {
        std::uint64_t v = (std::uint64_t)a.d[i + w] >> b;
        if (b && i + w + 1 < a.n)
            v |= (std::uint64_t)a.d[i + w + 1] << (32 - b);
        dst.d[i] = (std::uint32_t)v;
        dst.n = i + 1;
    }
    while (dst.n > 0 && dst.d[dst.n - 1] == 0)
        --dst.n;
}
BSPAI_HD inline int big_cmp(const Big &a, const Big &b) {
    if (a.n != b.n)
        return a.n < b.n ? -1 : 1;
    for (int i = a.n - 1; i >= 0; --i)
        if (a.d[i] != b.d[i])
            return a.d[i] < b.d[i] ? -1 : 1;
    return 0;
}
BSPAI_HD inline void big_sub(Big &a, const Big &b) { // a -= b, a >= b
    std::int64_t borrow = 0;
    for (int i = 0; i < a.n; ++i) {
        std::int64_t t = (std::int64_t)a.d[i] - (i < b.n ? (std::int64_t)b.d[i] : 0) - borrow;
        borrow = t < 0;
        if (t < 0)
            t += (std::int64_t)1 << 32;
        a.d[i] = (std::uint32_t)t;
    }
    while (a.n > 0 && a.d[a.n - 1] == 0)
        --a.n;
}
BSPAI_HD inline int big_bit(const Big &a, int i) {
    const int w = i / 32;
    return w < a.n ? (int)((a.d[w] >> (i % 32)) & 1u) : 0;
}
BSPAI_HD inline void big_pow5_mul(Big &a, int k) { // a *= 5^k
    while (k >= 13) {
        big_mul_small(a, 1220703125u); // 5^13
        k -= 13;
    }
    std::uint32_t m = 1;
    while (k-- > 0)
        m *= 5;
    if (m > 1)
        big_mul_small(a, m);
}

// Digits of a decimal token: the significant digits (leading zeros gone) are
// the characters of [db, de) other than '.', and value = D * 10^q with D the
// integer they form.
struct DecimalDigits {
    const char *db, *de;
    int nd;     // significant digits
    long long q;
};

// Exact conversion of D * 10^q (D != 0).  At most kMaxDigits digits form D;
// further digits only mark the value as above D * 10^q' (sticky).
constexpr int kMaxDigits = 800;
BSPAI_HD inline double mm_exact_decimal(bool neg, const DecimalDigits &dd, bool &range) {
    Big D, R, B;
    big_set(D, 0);
    int used = 0;
    bool sticky = false;
    long long q = dd.q;
    for (const char *p = dd.db; p < dd.de; ++p) {
        if (*p == '.')
            continue;
        if (used < kMaxDigits) {
            big_mul_small(D, 10u, (std::uint32_t)(*p - '0'));
            ++used;
        } else {
            sticky = sticky || *p != '0';
            ++q; // a digit dropped: the kept ones count ten times more
        }
    }
    // leading exponent: value in [10^(lead), 10^(lead+1))
    const long long lead = q + used - 1;
    if (lead > 309) {
        range = true;
        return inf_of(neg);
    }
    if (lead < -326) {
        range = true;
        return make_double(neg, 0, 0);
    }
    if (q >= 0) {
        // value = D * 5^q * 2^q: the top 64 bits of N = D 5^q and the rest.
        big_pow5_mul(D, (int)q);
        const int bits = big_bits(D);
        std::uint64_t top = 0;
        bool rest = sticky;
        const int take = bits < 64 ? bits : 64;
        for (int i = 0; i < take; ++i)
            top = (top << 1) | (std::uint64_t)big_bit(D, bits - 1 - i);
        for (int i = 0; i < bits - take && !rest; ++i)
            rest = big_bit(D, i) != 0;
        return round_binary(neg, top, (int)q + (bits - take), rest, range);
    }
    // q < 0: value = D / (5^|q| 2^|q|).  The 64-bit quotient
    // floor(D 2^t / (5^|q| 2^|q|)) and whether a remainder is left.
    const int k = (int)(-q);
    big_set(B, 1);
    big_pow5_mul(B, k);
    // choose t with the quotient of D 2^t / 5^k in [2^62, 2^64)
    const int t = 63 - (big_bits(D) - big_bits(B));
    // value = (D 2^t / 5^k) * 2^(-t-k)
    Big N = D;
    if (t > 0)
        big_shl(N, t);
    else if (t < 0)
        big_shl(B, -t); // D / (5^k 2^-t): the divisor absorbs the shift
    // restoring division: N >> 64 < B, then the low 64 bits of N one by one
    big_shr(R, N, 64);
    std::uint64_t quo = 0;
    for (int i = 63; i >= 0; --i) {
        big_shl(R, 1);
        if (big_bit(N, i)) {
            if (R.n == 0)
                R.n = 1, R.d[0] = 0;
            R.d[0] |= 1u;
        }
        quo <<= 1;
        if (big_cmp(R, B) >= 0) {
            big_sub(R, B);
            quo |= 1;
        }
    }
    // value = D / (5^k 2^k) = (D 2^t / 5^k) 2^(-t-k) = (quo + R / B) 2^(-t-k)
    return round_binary(neg, quo, -t - k, sticky || R.n != 0, range);
}

// Decimal conversion: returns false (kNeedsExact) where the exact method is
// required; else the correctly rounded value and its range flag.
BSPAI_HD inline bool mm_fast_decimal(bool neg, const DecimalDigits &dd, double &v, bool &range) {
    range = false;
    if (dd.nd > 19)
        return false;
    std::uint64_t w = 0;
    for (const char *p = dd.db; p < dd.de; ++p)
        if (*p != '.')
            w = w * 10 + (std::uint64_t)(*p - '0');
    const long long q = dd.q;
    // Clinger: w and 10^|q| exact doubles, one rounding
    constexpr double p10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    if (w <= (1ull << 53) && q >= -22 && q <= 22) {
        const double x = (double)w;
        const double r = q >= 0 ? x * p10[q] : x / p10[-q];
        v = neg ? -r : r;
        return true;
    }
    if (q < kPow5Min || q > kPow5Max)
        return false;
    // Eisel-Lemire: w 10^q = (w 2^lz) * 5^q * 2^(q - lz), 5^q ~ T 2^(g - 127)
    const int lz = clz64(w);
    const std::uint64_t wn = w << lz;
    const std::uint64_t *T = kPow5Table[q - kPow5Min];
    std::uint64_t h1, l1, h2, l2;
    mul64(wn, T[0], h1, l1);
    mul64(wn, T[1], h2, l2);
    // top 128 bits of the 192-bit product: (h1, l1) + h2 (carry into h1)
    std::uint64_t H = h1, L = l1 + h2;
    if (L < h2)
        ++H;
    const int upper = (int)(H >> 63); // H in [2^62, 2^64)
    const int shift = upper + 9;      // keeps 54 bits: 53 and a rounding bit
    std::uint64_t m = H >> shift;
    const std::uint64_t mask = (1ull << shift) - 1;
    const std::uint64_t low = H & mask;
    // The computed top 128 bits are within two units of L of the exact ones
    // (T is within one unit of 5^q's scaled value, the lower product is
    // truncated once): near a zero or a carrying sticky part, decide exactly.
    if ((low == 0 && L <= 2) || (low == mask && L >= ~0ull - 2))
        return false;
    // w 5^q = P 2^(g - 127 - lz) with P = wn T ~ H 2^128, so
    // value ~ H 2^(1 + g - lz + q) = m 2^E, E = shift + 1 + g - lz + q
    const int g = kPow5Log2[q - kPow5Min];
    const int E = shift + 1 + g - lz + (int)q;
    const int ue = E + 53; // unbiased exponent of the rounded value before a carry
    // subnormal, near the bottom (tininess must be decided exactly) or near
    // overflow: exact method
    if (ue <= -1021 || ue >= 1023)
        return false;
    // round half even on m's last bit; the sticky part is nonzero here
    // (low == 0 && L <= 2 was sent to the exact method)
    const std::uint64_t lastbit = m & 1;
    m >>= 1;
    m += lastbit; // rounding bit set and sticky nonzero: up; else down
    int ex = ue;
    if (m == (1ull << 53)) {
        m >>= 1;
        ++ex;
    }
    v = make_double(neg, (std::uint64_t)(ex + 1023), m - (1ull << 52));
    return true;
}

// Hexadecimal significand digits [hb, he) (with one optional '.'), binary
// exponent pexp: value = digits * 2^(pexp - 4 * fraction digits).
BSPAI_HD inline double mm_hex(bool neg, const char *hb, const char *he, long long pexp,
                              bool &range) {
    range = false;
    std::uint64_t mant = 0;
    int kept = 0;
    bool sticky = false, after = false;
    long long scale = pexp;
    bool lead = true;
    for (const char *p = hb; p < he; ++p) {
        if (*p == '.') {
            after = true;
            continue;
        }
        const char c = *p;
        const int dv = c <= '9' ? c - '0' : (c | 0x20) - 'a' + 10;
        if (lead && dv == 0) {
            if (after)
                scale -= 4;
            continue;
        }
        lead = false;
        if (kept < 15) { // 60 bits: room for the shift below
            mant = (mant << 4) | (std::uint64_t)dv;
            ++kept;
            if (after)
                scale -= 4;
        } else {
            sticky = sticky || dv != 0;
            if (!after)
                scale += 4;
        }
    }
    if (mant == 0)
        return make_double(neg, 0, 0);
    if (scale > 4000)
        scale = 4000;
    if (scale < -4000)
        scale = -4000;
    // glibc keeps a hexadecimal significand of more than 53 bits as 53 bits,
    // a rounding bit (the 54th) and a sticky flag, and for a subnormal result
    // shifts the 53 bits right without the 54th bit: that bit counts neither
    // for the rounding nor for the range error.  Kept here, so that the same
    // tokens are accepted and give the same values.
    const int len = 64 - clz64(mant);
    if (len >= 54 && (int)scale + len - 1 < -1022)
        mant &= ~(1ull << (len - 54));
    return round_binary(neg, mant, (int)scale, sticky, range);
}

BSPAI_HD inline char lower(char c) { return (c >= 'A' && c <= 'Z') ? (char)(c | 0x20) : c; }
BSPAI_HD inline bool match_ci(const char *p, const char *e, const char *word) {
    for (; *word; ++word, ++p)
        if (p >= e || lower(*p) != *word)
            return false;
    return true;
}
BSPAI_HD inline bool is_digit(char c) { return c >= '0' && c <= '9'; }
BSPAI_HD inline bool is_hex(char c) {
    return is_digit(c) || (lower(c) >= 'a' && lower(c) <= 'f');
}

// Decimal exponent digits after a marker: saturating value; returns the end.
BSPAI_HD inline const char *read_exponent(const char *p, const char *e, long long &ex) {
    // p points at the marker (e/E or p/P); consumed only with at least one digit
    const char *s = p + 1;
    bool eneg = false;
    if (s < e && (*s == '+' || *s == '-')) {
        eneg = *s == '-';
        ++s;
    }
    if (s >= e || !is_digit(*s)) {
        ex = 0;
        return p;
    }
    long long x = 0;
    while (s < e && is_digit(*s)) {
        if (x < 100000000)
            x = x * 10 + (*s - '0');
        ++s;
    }
    ex = eneg ? -x : x;
    return s;
}

// The value token [b, e) (no white space inside).  Returns the LineCode of
// the value (kLineOk, kBadValue, kNonFinite, kIntegerField, or kNeedsExact
// when kExact is false and the exact conversion is needed); v receives the
// value.  (kExact = false leaves the big-integer conversion out of the code,
// and with it its 1.6 KB of private memory per thread.)
template <bool kExact>
BSPAI_HD inline int parse_value(const char *b, const char *e, bool integer_field, double &v) {
    const char *p = b;
    bool neg = false;
    if (p < e && (*p == '+' || *p == '-')) {
        neg = *p == '-';
        ++p;
    }
    const char *end = b; // end of the converted prefix
    bool range = false;
    bool finite = true;
    if (match_ci(p, e, "inf")) {
        end = match_ci(p, e, "infinity") ? p + 8 : p + 3;
        v = inf_of(neg);
        finite = false;
    } else if (match_ci(p, e, "nan")) {
        end = p + 3;
        if (end < e && *end == '(') {
            const char *s = end + 1;
            while (s < e && (is_digit(*s) || (lower(*s) >= 'a' && lower(*s) <= 'z') || *s == '_'))
                ++s;
            if (s < e && *s == ')')
                end = s + 1;
        }
        v = nan_of(neg);
        finite = false;
    } else if (p + 1 < e && p[0] == '0' && (p[1] == 'x' || p[1] == 'X') &&
               ((p + 2 < e && is_hex(p[2])) ||
                (p + 3 < e && p[2] == '.' && is_hex(p[3])))) {
        const char *s = p + 2, *hb = s;
        while (s < e && is_hex(*s))
            ++s;
        if (s < e && *s == '.') {
            ++s;
            while (s < e && is_hex(*s))
                ++s;
        }
        const char *he = s;
        long long pexp = 0;
        if (s < e && (*s == 'p' || *s == 'P'))
            s = read_exponent(s, e, pexp);
        end = s;
        v = mm_hex(neg, hb, he, pexp, range);
    } else {
        // decimal
        const char *s = p;
        const char *ib = s;
        while (s < e && is_digit(*s))
            ++s;
        const char *ie = s, *fb = s, *fe = s;
        if (s < e && *s == '.') {
            fb = s + 1;
            fe = fb;
            while (fe < e && is_digit(*fe))
                ++fe;
            s = fe;
        }
        const int nint = (int)(ie - ib), nfrac = (int)(fe - fb);
        if (nint + nfrac == 0) {
            end = b; // nothing converted
        } else {
            long long ex = 0;
            if (s < e && (*s == 'e' || *s == 'E'))
                s = read_exponent(s, e, ex);
            end = s;
            // value = (all digits as one integer) * 10^(ex - nfrac).  The
            // significant span [db, de): leading zeros skipped (through the
            // point when the integer part is all zeros), trailing zeros
            // dropped (each one multiplies the rest by ten: q + 1).
            long long q = ex - nfrac;
            const char *db = ib;
            while (db < ie && *db == '0')
                ++db;
            if (db == ie) {
                db = fb;
                while (db < fe && *db == '0')
                    ++db;
            }
            const char *de = nfrac ? fe : ie;
            if (db >= de) {
                de = db; // no nonzero digit
            } else {
                while (de > db && (de[-1] == '0' || de[-1] == '.')) {
                    if (de[-1] == '0')
                        ++q;
                    --de;
                }
            }
            int nd = 0;
            for (const char *t = db; t < de; ++t)
                nd += *t != '.';
            if (nd == 0) {
                v = make_double(neg, 0, 0);
            } else {
                DecimalDigits dd{db, de, nd, q};
                if (!mm_fast_decimal(neg, dd, v, range)) {
                    if constexpr (!kExact)
                        return kNeedsExact;
                    else
                        v = mm_exact_decimal(neg, dd, range);
                }
            }
        }
    }
    if (end == b || range)
        return kBadValue;
    if (end != e || !finite)
        return kNonFinite;
    if (integer_field)
        for (const char *t = b; t < e; ++t)
            if (!(is_digit(*t) || *t == '+' || *t == '-'))
                return kIntegerField;
    return kLineOk;
}

// A data line [b, e) of a file with `rows` x `cols` and the value field kind.
// Returns its LineCode and, for kLineOk, the zero-based row, column and value.
template <bool kExact>
BSPAI_HD inline int parse_entry_line(const char *b, const char *e, std::uint64_t rows,
                                     std::uint64_t cols, bool integer_field, std::uint64_t &r,
                                     std::uint64_t &c, double &v) {
    const char *fb[3], *fe[3];
    const int nf = split_fields(b, e, fb, fe, 3);
    if (nf < 3)
        return kMalformedEntry;
    if (nf > 3)
        return kExtraField;
    std::uint64_t r1 = 0, c1 = 0;
    int code = parse_index(fb[0], fe[0], r1);
    if (code)
        return code;
    code = parse_index(fb[1], fe[1], c1);
    if (code)
        return code;
    if (r1 == 0 || c1 == 0 || r1 > rows || c1 > cols)
        return kIndexRange;
    code = parse_value<kExact>(fb[2], fe[2], integer_field, v);
    if (code)
        return code;
    r = r1 - 1;
    c = c1 - 1;
    return kLineOk;
}

} // namespace bspai_mm
