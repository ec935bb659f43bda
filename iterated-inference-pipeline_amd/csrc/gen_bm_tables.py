#!/usr/bin/env python3
"""Generate the lookup tables of the FP64 Box-Muller transform (plom_bm_tables.h).

The transform works directly on the two 32-bit integers x, y of a pair
(u1 = (x + 1/2) 2^-32, u2 = (y + 1/2) 2^-32):

* log u1: u1 = 2^e m with m in [1, 2); k = the top 7 fraction bits of m; c_k = 1 +
  (k + 1/2)/128 and invc_k = 1/c_k rounded to double.  Then
  log u1 = e ln2 + L_k + log1p(z) with z = m invc_k - 1 (one fma, |z| < 2^-8) and
  L_k = -log(invc_k) (so the rounding of invc_k cancels).  For k >= 64 (m >= 1.5) the
  table holds L_k - ln2 and the exponent is taken as e + 1: then for u1 close to 1 the
  result is a small table value plus the polynomial, with no cancellation against ln 2,
  and sqrt(-2 log u1) keeps its absolute accuracy there.  The last entry uses c = 2
  (invc = 1/2, table value 0, both exact): the interval that ends at u1 = 1 then gives
  log u1 = log1p(m/2 - 1) with z exact, so log u1 keeps full relative accuracy as u1 -> 1
  (|z| <= 2^-8 still holds there).
* sin/cos(2 pi u2): a_k = 2 pi (k + 1/2)/256 for k = the top 8 bits of y; the rest of
  the angle, d = 2 pi (r + 1/2) 2^-32 with r = (y mod 2^24) - 2^23, is at most pi/256 in
  magnitude; sin/cos(a_k + d) from sin/cos a_k and short Taylor polynomials of d.

Layout (PLOM_BM_TAB_SIZE doubles): [2k] = invc_k, [2k+1] = L_k for k < 128, then
[256 + 2k] = sin a_k, [256 + 2k + 1] = cos a_k for k < 256.  Values are computed with
mpmath at 60 digits and rounded to nearest.

Run:  python gen_bm_tables.py > plom_bm_tables.h
"""
import mpmath as mp

mp.mp.dps = 60

LOG_BITS = 7
SC_BITS = 8


def rn(x):
    """mpmath value rounded to the nearest double."""
    return float(mp.mpf(x))


def hexd(x):
    return float.hex(x)


def main():
    vals = []
    ln2 = mp.log(2)
    for k in range(1 << LOG_BITS):
        c = 1 + (mp.mpf(k) + mp.mpf(1) / 2) / (1 << LOG_BITS)
        if k == (1 << LOG_BITS) - 1:
            c = mp.mpf(2)
        invc = rn(1 / c)
        L = -mp.log(mp.mpf(invc))
        if k >= (1 << (LOG_BITS - 1)):
            L = L - ln2
        vals += [invc, rn(L)]
    for k in range(1 << SC_BITS):
        a = 2 * mp.pi * (mp.mpf(k) + mp.mpf(1) / 2) / (1 << SC_BITS)
        vals += [rn(mp.sin(a)), rn(mp.cos(a))]
    print("/* Generated by gen_bm_tables.py: lookup tables of the FP64 Box-Muller transform. Do not edit. */")
    print("#ifndef PLOM_BM_TABLES_H")
    print("#define PLOM_BM_TABLES_H")
    print("#define PLOM_BM_LOG_BITS %d" % LOG_BITS)
    print("#define PLOM_BM_SC_BITS %d" % SC_BITS)
    print("#define PLOM_BM_SC_OFFSET %d" % (2 << LOG_BITS))
    print("#define PLOM_BM_TAB_SIZE %d" % len(vals))
    print("#define PLOM_BM_TAB_VALUES { \\")
    for i in range(0, len(vals), 4):
        print("    " + ", ".join(hexd(v) for v in vals[i:i + 4]) + ("," if i + 4 < len(vals) else "") + " \\")
    print("}")
    print("#endif")


if __name__ == "__main__":
    main()
