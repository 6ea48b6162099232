// mm_parse.h -- parsing of MatrixMarket entry lines: tokens, decimal
// integers and correctly rounded decimal -> binary64 conversion.
//
// Shared by the device kernels in mm.hip and by a host build that the CPU
// tests compile against strtod.  The includer defines SABLE_HD (the function
// qualifier: `__host__ __device__` in HIP code) before including this file.
//
// The reference reads every token with JavaScript's Number()
// (seq/src/sparse32.js:1916-1946), which rounds a decimal literal to the
// nearest binary64, ties to even.  The conversion here gives the same bits:
//
// * Fast path: the Eisel-Lemire method (D. Lemire, "Number Parsing at a
//   Gigabyte per Second", Software: Practice and Experience 51(8), 2021).
//   The first 19 significant digits w and the decimal exponent q give
//   v = w * 10^q = (w << lz) * T(q) * 2^(g(q) + q - lz), where T(q) is a
//   128-bit normalised power of five (pow5_table.inc).  The 192-bit product
//   P = (w << lz) * T(q) is exact for 0 <= q <= 55.  Otherwise |P - true| < 2^64
//   with a known sign (T is truncated for q > 55, rounded up for q < 0), so
//   the top 64 bits of P are the true top bits, and the bits below are
//   nonzero, unless the middle 64 bits of P are all ones (q > 55) or all
//   zeros (q < 0).  Those cases are reported as ambiguous.
// * More than 19 significant digits: the value lies strictly between the
//   conversions of w and w + 1; when both agree that is the answer.
// * Every other case goes to exact_bits(): big-integer comparison of the
//   decimal against the midpoints next to a candidate (up to 800
//   significant digits; further digits only act as a sticky bit, which is
//   exact because a binary64 midpoint has at most 767 significant digits).
#pragma once
#include <stdint.h>

#ifndef SABLE_HD
#error "define SABLE_HD before including mm_parse.h"
#endif

namespace sable {
namespace mm {

struct Pow5 {
    uint64_t hi, lo;
    int32_t g;
};

constexpr uint64_t INF_BITS = 0x7ff0000000000000ull;
constexpr uint64_t FRAC_MASK = (1ull << 52) - 1;

// status codes of a parsed line
enum : int {
    LINE_BLANK = 0,
    LINE_ENTRY = 1,
    E_ROW_TOKEN = -1,   // row index is not a decimal integer
    E_COL_TOKEN = -2,   // column index missing or not a decimal integer
    E_VAL_TOKEN = -3,   // value missing or not a decimal number
    E_RANGE = -4,       // index outside [1, nrows] x [1, ncols]
    E_EXACT = -5,       // value needs more precision than exact_bits holds
};

SABLE_HD inline bool is_blank(int c) { return c == ' ' || c == '\t' || c == '\r'; }
SABLE_HD inline bool is_delim(int c) { return c < 0 || c == '\n' || is_blank(c); }

template <class Src>
SABLE_HD inline int64_t skip_blanks(const Src& s, int64_t p) {
    while (is_blank(s(p))) ++p;
    return p;
}

// Decimal integer token ([+]digits) at p; *end is set past it.  Returns -1
// when the token is not an unsigned integer, -2 when it exceeds 2^62.
template <class Src>
SABLE_HD inline int64_t parse_index(const Src& s, int64_t p, int64_t* end) {
    int c = s(p);
    if (c == '+') c = s(++p);
    if (c < '0' || c > '9') return -1;
    int64_t v = 0;
    bool big = false;
    while (c >= '0' && c <= '9') {
        if (v > (int64_t(1) << 58)) big = true;
        else v = v * 10 + (c - '0');
        c = s(++p);
    }
    *end = p;
    if (!is_delim(c)) return -1;
    return big ? -2 : v;
}

struct Dec {
    uint64_t w;   // first <= 19 significant digits
    int64_t q;    // v ~= w * 10^q
    bool trunc;   // a nonzero digit was dropped after the 19th
    bool neg;
    int kind;     // 0 finite, 1 infinity, -1 not a number literal
};

// JavaScript StringNumericLiteral for decimals:
//   [+-] ( "Infinity" | digits [ "." [digits] ] | "." digits ) [ (e|E) [+-] digits ]
template <class Src>
SABLE_HD inline Dec scan_decimal(const Src& s, int64_t p, int64_t* end) {
    Dec d{0, 0, false, false, -1};
    int c = s(p);
    if (c == '+' || c == '-') {
        d.neg = (c == '-');
        c = s(++p);
    }
    if (c == 'I') {
        const char* inf = "Infinity";
        for (int k = 0; k < 8; ++k)
            if (s(p + k) != inf[k]) return d;
        p += 8;
        *end = p;
        if (!is_delim(s(p))) return d;
        d.kind = 1;
        return d;
    }
    int ndig = 0;         // significant digits seen
    int64_t frac = 0;     // digits after the point (all of them)
    int64_t dropped = 0;  // significant digits beyond the 19th
    bool any = false;
    bool point = false;
    for (;; c = s(++p)) {
        if (c == '.') {
            if (point) break;
            point = true;
            continue;
        }
        if (c < '0' || c > '9') break;
        any = true;
        if (point) ++frac;
        if (ndig == 0 && c == '0') continue;  // leading zero
        if (ndig < 19) {
            d.w = d.w * 10 + (uint64_t)(c - '0');
            ++ndig;
        } else {
            ++dropped;
            if (c != '0') d.trunc = true;
        }
    }
    if (!any) return d;
    int64_t ex = 0;
    if (c == 'e' || c == 'E') {
        c = s(++p);
        bool eneg = false;
        if (c == '+' || c == '-') {
            eneg = (c == '-');
            c = s(++p);
        }
        if (c < '0' || c > '9') return d;
        for (; c >= '0' && c <= '9'; c = s(++p))
            if (ex < 100000000) ex = ex * 10 + (c - '0');
        if (eneg) ex = -ex;
    }
    *end = p;
    if (!is_delim(c)) return d;
    d.q = ex - frac + dropped;
    d.kind = 0;
    return d;
}

SABLE_HD inline uint64_t umulhi64(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

// Eisel-Lemire: bits of the positive binary64 nearest w * 10^q.  Returns
// false when the 192-bit product cannot decide the rounding; *out then
// holds a candidate within one unit of the answer.
SABLE_HD inline bool el_bits(uint64_t w, int64_t q, const Pow5* tab, uint64_t* out) {
    if (w == 0 || q < -342) {  // w*10^q < 2^64 * 10^-343 < 2^-1075: rounds to 0
        *out = 0;
        return true;
    }
    if (q > 308) {  // >= 10^309
        *out = INF_BITS;
        return true;
    }
    const int lz = __builtin_clzll(w);
    const uint64_t W = w << lz;
    const Pow5 t = tab[q + 342];
    // P = W * (hi:lo) as three 64-bit words P2:P1:P0
    const uint64_t a_lo = W * t.lo, a_hi = umulhi64(W, t.lo);
    const uint64_t b_lo = W * t.hi, b_hi = umulhi64(W, t.hi);
    const uint64_t P0 = a_lo;
    const uint64_t P1 = a_hi + b_lo;
    const uint64_t P2 = b_hi + (P1 < a_hi ? 1 : 0);
    const bool exact = (q >= 0 && q <= 55);
    // ambiguous: a carry or borrow out of the low 128 bits is possible; the
    // result below is then only a candidate
    const bool ok = exact || (q >= 0 ? (P1 != ~0ull) : (P1 != 0));
    const int top = (int)(P2 >> 63) + 190;          // leading bit of P
    const int64_t E = top + (int64_t)t.g + q - lz;  // v in [2^E, 2^(E+1))
    if (E > 1023) {
        *out = INF_BITS;
        return ok;
    }
    const int64_t k = (E >= -1022) ? 53 : E + 1075;  // significand bits kept
    if (k < 0) {
        *out = 0;
        return ok;
    }
    const int sh = top - 128 - (int)k;  // P2 bits below the rounding bit, in [9, 63]
    const uint64_t mr = P2 >> sh;
    const bool sticky = !exact || (P2 & ((1ull << sh) - 1)) != 0 || P1 != 0 || P0 != 0;
    uint64_t mant = mr >> 1;
    if ((mr & 1) && (sticky || (mant & 1))) ++mant;
    if (k == 53) {
        int64_t e = E;
        if (mant == (1ull << 53)) {
            mant >>= 1;
            ++e;
        }
        if (e > 1023) {
            *out = INF_BITS;
            return ok;
        }
        *out = ((uint64_t)(e + 1023) << 52) | (mant & FRAC_MASK);
    } else {
        *out = mant;  // subnormal; 2^52 here is the smallest normal
    }
    return ok;
}

// w * 10^q for w < 2^64.  Trailing zeros of w move into q first.  Then, when
// w <= 2^53 and |q| <= 22, both w and 10^|q| are exact binary64 values and one
// IEEE multiply or divide rounds correctly (Clinger's fast path; W. D.
// Clinger, "How to Read Floating Point Numbers Accurately", PLDI 1990).  This
// also covers the short exact binary fractions (0.5, 0.25, 1.5 ...) whose
// Eisel-Lemire product is ambiguous by construction.  Everything else goes
// through el_bits().
SABLE_HD inline bool conv_bits(uint64_t w, int64_t q, const Pow5* tab, uint64_t* out) {
    if (w == 0) {
        *out = 0;
        return true;
    }
    while (w % 10 == 0) {
        w /= 10;
        ++q;
    }
    if (w <= (1ull << 53) && q >= -22 && q <= 22) {
        const double p10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
        const double x = (double)w;
        const double r = q >= 0 ? x * p10[q] : x / p10[-q];
        *out = __builtin_bit_cast(uint64_t, r);
        return true;
    }
    return el_bits(w, q, tab, out);
}

// Fast conversion of a scanned token: *ambiguous is set when exact_bits()
// must decide; the result then is a candidate within one unit of the answer.
SABLE_HD inline uint64_t fast_bits(const Dec& d, const Pow5* tab, bool* ambiguous) {
    *ambiguous = false;
    uint64_t b;
    if (d.kind == 1) {
        b = INF_BITS;
    } else if (!d.trunc) {
        *ambiguous = !conv_bits(d.w, d.q, tab, &b);
    } else {
        uint64_t b2 = 0;
        const bool ok1 = conv_bits(d.w, d.q, tab, &b);
        const bool ok2 = conv_bits(d.w + 1, d.q, tab, &b2);
        *ambiguous = !(ok1 && ok2 && b == b2);
    }
    return b | (d.neg ? (1ull << 63) : 0);
}

// ---------------------------------------------------------------- exact path
constexpr int BIG_LIMBS = 136;  // 4352 bits
constexpr int MAX_SIG_DIGITS = 800;

struct Big {
    uint32_t d[BIG_LIMBS];
    int n;
    bool overflow;
};

SABLE_HD inline void big_set(Big& a, uint64_t v) {
    a.n = 0;
    a.overflow = false;
    while (v) {
        a.d[a.n++] = (uint32_t)v;
        v >>= 32;
    }
}

SABLE_HD inline void big_copy(Big& a, const Big& b) {
    a.n = b.n;
    a.overflow = b.overflow;
    for (int i = 0; i < b.n; ++i) a.d[i] = b.d[i];
}

SABLE_HD inline void big_mul_add(Big& a, uint32_t m, uint32_t add) {
    uint64_t carry = add;
    for (int i = 0; i < a.n; ++i) {
        const uint64_t t = (uint64_t)a.d[i] * m + carry;
        a.d[i] = (uint32_t)t;
        carry = t >> 32;
    }
    if (carry) {
        if (a.n < BIG_LIMBS) a.d[a.n++] = (uint32_t)carry;
        else a.overflow = true;
    }
}

SABLE_HD inline void big_mul_pow5(Big& a, int64_t p) {
    while (p >= 13) {
        big_mul_add(a, 1220703125u, 0);  // 5^13
        p -= 13;
    }
    uint32_t m = 1;
    while (p-- > 0) m *= 5;
    if (m != 1) big_mul_add(a, m, 0);
}

SABLE_HD inline void big_shl(Big& a, int64_t bits) {
    if (a.n == 0 || bits <= 0) return;
    const int64_t limbs = bits >> 5;
    const int r = (int)(bits & 31);
    const int64_t nn = a.n + limbs + 1;
    if (nn > BIG_LIMBS) {
        a.overflow = true;
        return;
    }
    a.d[nn - 1] = 0;
    for (int64_t i = a.n - 1; i >= 0; --i) {
        const uint64_t v = (uint64_t)a.d[i] << r;
        a.d[i + limbs + 1] |= (uint32_t)(v >> 32);
        a.d[i + limbs] = (uint32_t)v;
    }
    for (int64_t i = 0; i < limbs; ++i) a.d[i] = 0;
    a.n = (int)nn;
    while (a.n > 0 && a.d[a.n - 1] == 0) --a.n;
}

SABLE_HD inline int big_cmp(const Big& a, const Big& b) {
    if (a.n != b.n) return a.n < b.n ? -1 : 1;
    for (int i = a.n - 1; i >= 0; --i)
        if (a.d[i] != b.d[i]) return a.d[i] < b.d[i] ? -1 : 1;
    return 0;
}

// Compares D * 10^q (+ a sticky fraction) with odd * 2^e2.
SABLE_HD inline int cmp_decimal(const Big& D, bool sticky, int64_t q, uint64_t odd, int64_t e2, Big& A, Big& B,
                                bool* overflow) {
    big_copy(A, D);
    big_set(B, odd);
    if (q >= 0) big_mul_pow5(A, q);
    else big_mul_pow5(B, -q);
    const int64_t k = q < e2 ? q : e2;
    big_shl(A, q - k);
    big_shl(B, e2 - k);
    if (A.overflow || B.overflow) *overflow = true;
    int c = big_cmp(A, B);
    if (c == 0 && sticky) c = 1;
    return c;
}

// odd * 2^e2 = the midpoint between the positive finite double with these
// bits and the next one up.
SABLE_HD inline void upper_mid(uint64_t bits, uint64_t* odd, int64_t* e2) {
    const uint64_t ef = bits >> 52, fr = bits & FRAC_MASK;
    const uint64_t M = ef ? (fr | (1ull << 52)) : fr;
    const int64_t e = ef ? (int64_t)ef - 1075 : -1074;
    *odd = 2 * M + 1;
    *e2 = e - 1;
}

// Exact conversion of the decimal token at p, starting from a candidate
// (positive bits) within a few units of the answer.  Returns the signed
// bits; *fail is set if the token holds more precision than Big can carry.
template <class Src>
SABLE_HD inline uint64_t exact_bits(const Src& s, int64_t p, uint64_t cand, Big* scratch3, bool* fail) {
    Big& D = scratch3[0];
    Big& A = scratch3[1];
    Big& B = scratch3[2];
    *fail = false;
    big_set(D, 0);
    int c = s(p);
    bool neg = false;
    if (c == '+' || c == '-') {
        neg = (c == '-');
        c = s(++p);
    }
    int ndig = 0;
    int64_t frac = 0, dropped = 0;
    bool sticky = false, point = false;
    uint32_t chunk = 0;
    int nchunk = 0;
    for (;; c = s(++p)) {
        if (c == '.') {
            if (point) break;
            point = true;
            continue;
        }
        if (c < '0' || c > '9') break;
        if (point) ++frac;
        if (ndig == 0 && c == '0') continue;
        if (ndig < MAX_SIG_DIGITS) {
            chunk = chunk * 10 + (uint32_t)(c - '0');
            ++ndig;
            if (++nchunk == 9) {
                big_mul_add(D, 1000000000u, chunk);
                chunk = 0;
                nchunk = 0;
            }
        } else {
            ++dropped;
            if (c != '0') sticky = true;
        }
    }
    if (nchunk) {
        uint32_t m = 1;
        for (int i = 0; i < nchunk; ++i) m *= 10;
        big_mul_add(D, m, chunk);
    }
    int64_t ex = 0;
    if (c == 'e' || c == 'E') {
        c = s(++p);
        bool eneg = false;
        if (c == '+' || c == '-') {
            eneg = (c == '-');
            c = s(++p);
        }
        for (; c >= '0' && c <= '9'; c = s(++p))
            if (ex < 100000000) ex = ex * 10 + (c - '0');
        if (eneg) ex = -ex;
    }
    const int64_t q = ex - frac + dropped;
    uint64_t b = cand & ~(1ull << 63);
    if (D.n == 0) return neg ? (1ull << 63) : 0;
    bool ovf = false;
    for (int guard = 0; guard < 8; ++guard) {
        uint64_t odd;
        int64_t e2;
        if (b < INF_BITS) {
            upper_mid(b, &odd, &e2);
            const int cu = cmp_decimal(D, sticky, q, odd, e2, A, B, &ovf);
            if (cu > 0 || (cu == 0 && (b & 1))) {
                ++b;
                continue;
            }
        }
        if (b > 0) {
            upper_mid(b - 1, &odd, &e2);
            const int cl = cmp_decimal(D, sticky, q, odd, e2, A, B, &ovf);
            if (cl < 0 || (cl == 0 && (b & 1))) {
                --b;
                continue;
            }
        }
        if (ovf) *fail = true;
        return b | (neg ? (1ull << 63) : 0);
    }
    *fail = true;
    return b | (neg ? (1ull << 63) : 0);
}

// ---------------------------------------------------------------- lines
struct Entry {
    int64_t row, col;
    int64_t vpos;  // byte position of the value token
    Dec v;
    int status;    // LINE_BLANK, LINE_ENTRY or an error code
};

// Parses one line starting at p: "row col [value ...]" separated by blanks.
// A line holding only blanks, or a comment ('%' after optional blanks), is
// LINE_BLANK.  Tokens after the ones the field needs are ignored, as the
// reference ignores them.
template <class Src>
SABLE_HD inline void parse_line(const Src& s, int64_t p, bool pattern, Entry* e) {
    e->status = LINE_BLANK;
    e->row = e->col = 0;
    e->vpos = -1;
    e->v = Dec{0, 0, false, false, 0};
    p = skip_blanks(s, p);
    int c = s(p);
    if (c < 0 || c == '\n' || c == '%') return;
    int64_t end = p;
    e->row = parse_index(s, p, &end);
    if (e->row < 0) {
        e->status = E_ROW_TOKEN;
        return;
    }
    p = skip_blanks(s, end);
    e->col = parse_index(s, p, &end);
    if (e->col < 0) {
        e->status = E_COL_TOKEN;
        return;
    }
    if (!pattern) {
        p = skip_blanks(s, end);
        e->vpos = p;
        e->v = scan_decimal(s, p, &end);
        if (e->v.kind < 0) {
            e->status = E_VAL_TOKEN;
            return;
        }
    }
    e->status = LINE_ENTRY;
}

}  // namespace mm
}  // namespace sable
