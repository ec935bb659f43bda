#!/usr/bin/env python3
"""Generate the Chebyshev coefficient tables behind plom_Phi (plom_phi_coeffs.h).

The standard normal CDF is evaluated piecewise:

* |x| <= 1:  Phi(x) = 1/2 + x * g(x^2),  g(u) = (Phi(sqrt u) - 1/2) / sqrt(u)
  on u in [0, 1];
* x < -1:    Phi(x) = phi(x) * M(t) with t = -x, the Mills ratio M(t) = Phi(-t)/phi(t).
  h(t) = t * M(t) is tabulated as a function of s = 1/t on the bands
  t in [1,2], [2,4], [4,8], [8,40];
* x > 1:     Phi(x) = 1 - Phi(-x).

Each piece is a Chebyshev interpolant (first-kind nodes, computed with mpmath at 60
digits) in a variable mapped to [-1, 1], evaluated by the Clenshaw recurrence.  The
degrees are chosen so that the truncated tail is below 1e-17 relative.

Run:  python gen_phi_table.py > plom_phi_coeffs.h
"""
import mpmath as mp

mp.mp.dps = 60


def cheb_interp(f, a, b, n):
    xs = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / (n + 1)) for k in range(n + 1)]
    fx = [f((b - a) / 2 * x + (a + b) / 2) for x in xs]
    c = []
    for j in range(n + 1):
        s = mp.fsum(fx[k] * mp.cos(mp.pi * j * (k + mp.mpf(1) / 2) / (n + 1)) for k in range(n + 1))
        c.append(2 * s / (n + 1))
    c[0] /= 2
    return c


def g_central(u):
    if u == 0:
        return 1 / mp.sqrt(2 * mp.pi)
    s = mp.sqrt(u)
    return (mp.ncdf(s) - mp.mpf(1) / 2) / s


def h_mills(s):
    t = 1 / s
    return t * mp.ncdf(-t) / mp.npdf(t)


PIECES = [
    # name, function, lo, hi (in the fitted variable), degree
    ("central", g_central, mp.mpf(0), mp.mpf(1), 14),
    ("mills_1_2", h_mills, mp.mpf(1) / 2, mp.mpf(1), 22),
    ("mills_2_4", h_mills, mp.mpf(1) / 4, mp.mpf(1) / 2, 20),
    ("mills_4_8", h_mills, mp.mpf(1) / 8, mp.mpf(1) / 4, 16),
    ("mills_8_40", h_mills, mp.mpf(1) / 40, mp.mpf(1) / 8, 18),
]


def main():
    out = []
    out.append("/* Generated by gen_phi_table.py: Chebyshev coefficients for plom_Phi. Do not edit. */")
    out.append("#ifndef PLOM_PHI_COEFFS_H")
    out.append("#define PLOM_PHI_COEFFS_H")
    for name, f, lo, hi, deg in PIECES:
        c = cheb_interp(f, lo, hi, deg)
        out.append("/* %s: variable in [%s, %s], degree %d, last coefficient %s */"
                   % (name, mp.nstr(lo, 6), mp.nstr(hi, 6), deg, mp.nstr(c[-1], 3)))
        out.append("#define PLOM_PHI_%s_LO %s" % (name.upper(), mp.nstr(lo, 20)))
        out.append("#define PLOM_PHI_%s_HI %s" % (name.upper(), mp.nstr(hi, 20)))
        out.append("#define PLOM_PHI_%s_DEG %d" % (name.upper(), deg))
        vals = ", ".join(float(x).hex() for x in c)
        out.append("#define PLOM_PHI_%s_COEFFS { %s }" % (name.upper(), vals))
    out.append("#endif")
    print("\n".join(out))


if __name__ == "__main__":
    main()
