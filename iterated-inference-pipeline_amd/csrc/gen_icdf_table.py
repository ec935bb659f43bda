#!/usr/bin/env python3
"""Generate the coefficient table of the FP64 Gaussian transform of the Euler noise
(plom_icdf_table.h): z = Phi^-1(u) on piecewise polynomials.

One 32-bit word w of a particle's stream gives one normal:
    v = w >> 1 (31 bits),  u = (v + 1/2) 2^-32 in [2^-33, 1/2),  z = +-Phi^-1(u),
the sign taken from the low bit of w (w & 1: negative Phi^-1(u) becomes positive), so z is
the inverse-CDF transform of the uniform (v + 1/2) 2^-32 on one half and its mirror image
on the other: z = Phi^-1(U) with U = (w' + 1/2) 2^-32 for a 32-bit w' that is w
re-arranged, exactly normal up to the 2^-32 quantisation of U.

Segments: u = 2^E m with E in [-33, -2] (32 binary octaves) and m in [1, 2); the top
SUB_BITS = 4 fraction bits of m pick one of 16 sub-segments of the octave, so the segment
index is the top bits of the double's high word, idx = (E + 33) 16 + s in [0, 512).  The
remaining fraction r in [0, 2^-4) gives the local coordinate t = 32 (1 + r) - 33 in
[-1, 1), exact in double, and z = sum_i c_i t^i (Horner, degree DEG = 6).

Coefficients: interpolation of Phi^-1 (mpmath erfinv at 50 digits) at the 7 Chebyshev
nodes of the segment, written in the monomial basis of t and rounded to double.  The
script prints the largest error of the double evaluation against mpmath over 65 points
of every segment (about 2e-14 absolute; the tests check it again).

Layout (coefficient-major, so that one LDS address and constant offsets reach all seven
coefficients of a segment): tab[i * NSEG + slot] = c_i of segment idx, where the slot is
idx rotated to the low bits of the converted word's exponent and top fraction bits, so
that the transform reads the slot straight from those bits (a shift and a mask):

  FP64: the word is converted as v1 = w | 1 = u 2^33, whose biased exponent is
        (E + 33) + 1023, so slot = (idx + 1023 * 16) mod NSEG = (idx + ROT64) mod NSEG;
  FP32: the top 24 bits are converted as v1 = (w >> 8) | 1 = u' 2^25, biased exponent
        (E + 25) + 127 = (E + 33) + 119, so slot = (idx + ROT32) mod NSEG.

The header also holds the coefficients rounded to float, for the FP32 stream, in the
FP32 slots.

Run:  python gen_icdf_table.py > plom_icdf_table.h
"""
import sys

import mpmath as mp
import numpy as np

mp.mp.dps = 50

SUB_BITS = 4
DEG = 6
E_MIN, E_MAX = -33, -2
NSEG = (E_MAX - E_MIN + 1) << SUB_BITS
ROT64 = (1023 << SUB_BITS) % NSEG   # slot of segment 0 (u in [2^-33, ...), v1 = 1.0)
ROT32 = (119 << SUB_BITS) % NSEG    # slot of segment 0 for the float conversion


def seg_u(idx, t):
    """u at local coordinate t of segment idx"""
    E = E_MIN + (idx >> SUB_BITS)
    s = idx & ((1 << SUB_BITS) - 1)
    return mp.ldexp(1 + (s + (t + 1) / 2) / mp.mpf(1 << SUB_BITS), E)


def phinv(u):
    return mp.sqrt(2) * mp.erfinv(2 * u - 1)


def fit(idx):
    n = DEG + 1
    nodes = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
    A = mp.matrix([[t ** i for i in range(n)] for t in nodes])
    c = mp.lu_solve(A, mp.matrix([phinv(seg_u(idx, t)) for t in nodes]))
    return [float(c[i]) for i in range(n)]


def horner(c, t):
    p = c[-1]
    for ci in reversed(c[:-1]):
        p = p * t + ci   # the table's users evaluate with fma; the error check below is looser
    return p


def main():
    coef = [fit(idx) for idx in range(NSEG)]
    worst = 0.0
    for idx in range(NSEG):
        for k in range(65):
            t = -1.0 + 2.0 * k / 64
            worst = max(worst, abs(horner(coef[idx], t) - float(phinv(seg_u(idx, mp.mpf(t))))))
    print("max |error| %.3g" % worst, file=sys.stderr)
    out = ["/* Generated by gen_icdf_table.py: piecewise polynomials of the inverse normal CDF on",
           " * [2^-33, 1/2) (see that script).  Do not edit. */",
           "#ifndef PLOM_ICDF_TABLE_H", "#define PLOM_ICDF_TABLE_H",
           "#define PLOM_ICDF_SUB_BITS %d" % SUB_BITS, "#define PLOM_ICDF_DEG %d" % DEG,
           "#define PLOM_ICDF_NSEG %d" % NSEG,
           "#define PLOM_ICDF_EB_MIN %d /* biased exponent of the first octave, 2^%d */" % (E_MIN + 1023, E_MIN),
           "#define PLOM_ICDF_TAB_SIZE %d" % ((DEG + 1) * NSEG),
           "#define PLOM_ICDF_ROT64 %d /* slot of segment 0 in the FP64 table */" % ROT64,
           "#define PLOM_ICDF_ROT32 %d /* slot of segment 0 in the FP32 table */" % ROT32,
           "#define PLOM_ICDF_TAB_VALUES { \\"]
    at64 = [(slot - ROT64) % NSEG for slot in range(NSEG)]   # segment held by each slot
    at32 = [(slot - ROT32) % NSEG for slot in range(NSEG)]
    for i in range(DEG + 1):
        for slot in range(0, NSEG, 4):
            out.append("    " + ", ".join(coef[at64[j]][i].hex() for j in range(slot, slot + 4)) + ", \\")
    out.append("}")
    out.append("/* the same coefficients rounded to float (the FP32 stream), in the FP32 slots */")
    out.append("#define PLOM_ICDF_TAB32_VALUES { \\")
    for i in range(DEG + 1):
        for slot in range(0, NSEG, 4):
            out.append("    " + ", ".join("%sf" % float(np.float32(coef[at32[j]][i])).hex() for j in range(slot, slot + 4)) + ", \\")
    out.append("}")
    out.append("#endif")
    print("\n".join(out))


if __name__ == "__main__":
    main()
