"""Writes pow5_table.inc: 128-bit normalised powers of five for the decimal
parser in mm_parse.h.

For every q in [-342, 308] the table holds T(q) and g(q) with
    5**q ~= T(q) * 2**g(q),   2**127 <= T(q) < 2**128.
* q >= 0: T is 5**q shifted to 128 bits, truncated (exact for q <= 55).
* q <  0: T is ceil(2**s / 5**-q) with s chosen so T has 128 bits, g = -s.
The rounding direction is what the parser's error analysis relies on:
the product w*T is below the true value for q > 55 and above it for q < 0.

Run: python gen_pow5.py [out]   (the committed file is this script's output)
"""
import os
import sys

QMIN, QMAX = -342, 308


def entry(q):
    if q >= 0:
        b = 5 ** q
        L = b.bit_length()
        t = b << (128 - L) if L <= 128 else b >> (L - 128)
        g = L - 128
    else:
        b = 5 ** (-q)
        L = b.bit_length()
        s = 127 + L
        t = -((-(1 << s)) // b)  # ceil
        g = -s
    assert (1 << 127) <= t < (1 << 128), q
    return t, g


def render():
    lines = ["// Generated by gen_pow5.py: 5^q ~= T * 2^g for q in [%d, %d]." % (QMIN, QMAX),
             "// Rows: {high 64 bits of T, low 64 bits of T, g}.",
             "#define SABLE_POW5_QMIN (%d)" % QMIN,
             "#define SABLE_POW5_QMAX (%d)" % QMAX]
    for q in range(QMIN, QMAX + 1):
        t, g = entry(q)
        lines.append("{0x%016xull, 0x%016xull, %d},  // %d" % (t >> 64, t & ((1 << 64) - 1), g, q))
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                             "pow5_table.inc")
    with open(out, "w") as f:
        f.write(render())
