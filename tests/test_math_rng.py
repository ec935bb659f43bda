"""Shared arithmetic (csrc/plom_math.h, csrc/plom_rng.h), host build, pinned against
published known answers and high-precision references."""
import math

import mpmath as mp
import numpy as np
import pytest
from scipy import special, stats

from oracle import oracle

mp.mp.dps = 40


def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32_10 (Salmon et al., SC'11)."""
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_xoshiro128ss_first_output_and_uniformity():
    out, _ = oracle.xoshiro([1, 2, 3, 4], 200000)
    # first output: rotl(s1*5, 7)*9 with s1 = 2 -> rotl(10, 7)*9 = 11520
    assert out[0] == 11520
    u = (np.asarray(out, dtype=np.float64) + 0.5) / 2 ** 32
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.01
    bits = np.unpackbits(np.asarray(out, dtype=np.uint32).view(np.uint8))
    assert abs(bits.mean() - 0.5) < 0.002


def xoshiro128ss_reference(s, n):
    """xoshiro128** 1.1 as published by Blackman and Vigna (prng.di.unimi.it, xoshiro128starstar.c):
    result = rotl(s1 * 5, 7) * 9; t = s1 << 9; s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3;
    s2 ^= t; s3 = rotl(s3, 11)."""
    M = 0xffffffff
    rotl = lambda x, k: ((x << k) | (x >> (32 - k))) & M
    s = list(s)
    out = []
    for _ in range(n):
        out.append(rotl((s[1] * 5) & M, 7) * 9 & M)
        t = (s[1] << 9) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 11)
    return out


def test_xoshiro128ss_stream_equals_the_published_algorithm():
    """The stream's words against the published step, from two seeds, over enough words to
    exercise every state word."""
    for seed in ([1, 2, 3, 4], [0x9e3779b9, 0x243f6a88, 0x85a308d3, 0xfffffffe]):
        out, _ = oracle.xoshiro(seed, 4000)
        assert list(np.asarray(out, dtype=np.uint64)) == xoshiro128ss_reference(seed, 4000)


def test_div_rc_equals_ieee_division():
    """x / c from the divisor's reciprocal with one correction step (plom_div_rc, the steps'
    division by a population size) against IEEE division: equal on random quotients with
    population-size and random divisors, on quotients whose mantissas lie next to 1 and 2
    (where q = RN(x RN(1/c)) is furthest from x / c), with the sign of x / c for signed
    zeros, and NaN for infinite x."""
    rng = np.random.default_rng(11)
    n = 400000
    x = np.exp(rng.uniform(np.log(1e-6), np.log(1e9), n))
    for c in (5e5, 1e6, 3e5, 123457.0, 1e6 / 3):
        np.testing.assert_array_equal(oracle.div_rc(x, c), x / c)
    c = np.exp(rng.uniform(-30, 30, n))
    np.testing.assert_array_equal(oracle.div_rc(x, c), x / c)
    xm = 2 - np.exp(rng.uniform(np.log(1e-16), np.log(1e-1), n))
    for cm in (2 - np.exp(rng.uniform(np.log(1e-16), np.log(1e-2), n)), 1 + np.exp(rng.uniform(np.log(1e-16), np.log(1e-2), n))):
        np.testing.assert_array_equal(oracle.div_rc(xm, cm), xm / cm)
    z = oracle.div_rc(np.array([0.0, -0.0, 0.0, -0.0]), np.array([5e5, 5e5, -5e5, -5e5]))
    assert list(np.signbit(z)) == [False, True, True, False] and np.all(z == 0.0)
    assert np.all(np.isnan(oracle.div_rc(np.array([np.inf, -np.inf]), 5e5)))


def test_normals_distribution():
    z = np.concatenate([oracle.normals(1, s, n, j, 64) for s in range(4) for n in range(16) for j in range(64)])
    assert stats.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1) < 0.03


def test_normals_are_a_function_of_the_key():
    a = oracle.normals(3, 4, 5, 6, 10)
    assert np.array_equal(a, oracle.normals(3, 4, 5, 6, 10))
    assert not np.array_equal(a, oracle.normals(3, 4, 5, 7, 10))
    assert not np.array_equal(a, oracle.normals(3, 4, 6, 6, 10))
    assert not np.array_equal(a, oracle.normals(3, 5, 5, 6, 10))


def ulp_err(got, ref):
    ref = float(ref)
    if ref == 0.0:
        return abs(got)
    return abs(got - ref) / math.ulp(ref)


def test_exp_log_sin_cos_accuracy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-700, 700, 400), rng.uniform(-1, 1, 200), rng.uniform(-50, 50, 200)])
    r = oracle.math(xs)
    for x, row in zip(xs, r):
        assert ulp_err(row[0], mp.exp(mp.mpf(x))) <= 2
        assert abs(row[3] - float(mp.sin(mp.mpf(x)))) <= 2 * 2 ** -53 * max(1.0, 1.0)
        assert abs(row[4] - float(mp.cos(mp.mpf(x)))) <= 2 * 2 ** -53
    pos = np.exp(rng.uniform(-740, 700, 600))
    pos = np.concatenate([pos, 1 + rng.uniform(-1e-6, 1e-6, 100), [1e-310, 5e-324]])
    r = oracle.math(pos)
    for x, row in zip(pos, r):
        assert ulp_err(row[1], mp.log(mp.mpf(x))) <= 2
        assert row[5] == math.sqrt(x)


def test_sqrt_rate():
    """The demographic-noise square root: IEEE sqrt on [2^-767, inf), 0 on [+0, 2^-767),
    NaN for negative numbers, -0, +inf and NaN."""
    rng = np.random.default_rng(3)
    big = np.concatenate([np.exp2(rng.uniform(-767, 1023, 20000)), rng.uniform(0, 1e6, 20000),
                          [2.0 ** -767, np.nextafter(2.0 ** -767, 1.0), 1.0, 4.0, np.finfo(float).max]])
    np.testing.assert_array_equal(oracle.math(big)[:, 7], np.sqrt(big))
    tiny = np.concatenate([[0.0, 5e-324, 2.0 ** -1022, np.nextafter(2.0 ** -767, 0.0)],
                           np.exp2(rng.uniform(-1074, -767.0001, 1000))])
    np.testing.assert_array_equal(oracle.math(tiny)[:, 7], 0.0)
    odd = np.array([-0.0, -1.0, -5e-324, np.inf, -np.inf, np.nan])
    assert np.all(np.isnan(oracle.math(odd)[:, 7]))


def test_special_values():
    r = oracle.math(np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -1.0, 1000.0, -1000.0]))
    assert r[0, 1] == -np.inf and r[2, 1] == np.inf and np.isnan(r[5, 1]) and np.isnan(r[4, 1])
    assert r[3, 0] == 0.0 and r[2, 0] == np.inf and r[6, 0] == np.inf and r[7, 0] == 0.0
    assert r[0, 2] == 0.5 and r[2, 2] == 1.0 and r[3, 2] == 0.0 and np.isnan(r[4, 2])


def test_phi_against_mpmath_and_scipy():
    xs = np.concatenate([np.linspace(-37.5, 8.0, 3001), [-1.0, 1.0, -2.0, 2.0, -4.0, 4.0, -8.0]])
    r = oracle.math(xs)[:, 2]
    for x, got in zip(xs, r):
        ref = float(mp.ncdf(mp.mpf(x)))
        assert abs(got - ref) <= 1e-14 * ref + 1e-300, (x, got, ref)
    np.testing.assert_allclose(r, special.ndtr(xs), rtol=3e-13, atol=0)


def test_phi_central_symmetry():
    xs = np.linspace(0, 6, 601)
    lo = oracle.math(-xs)[:, 2]
    hi = oracle.math(xs)[:, 2]
    np.testing.assert_allclose(lo + hi, 1.0, rtol=0, atol=2.3e-16)


def test_sincos_domain():
    """sin/cos are defined on |x| <= 2^20 and NaN outside (the same on host and device)."""
    r = oracle.math(np.array([2.0 ** 20, -2.0 ** 20, 2.0 ** 20 * 1.0000001, 1e300]))
    assert np.isfinite(r[:2, 3:5]).all() and np.isnan(r[2:, 3:5]).all()


def bm_exact(x, y):
    """Box-Muller of the words x, y in 40-digit arithmetic."""
    u1 = (mp.mpf(int(x)) + mp.mpf(1) / 2) / 2 ** 32
    u2 = (mp.mpf(int(y)) + mp.mpf(1) / 2) / 2 ** 32
    r = mp.sqrt(-2 * mp.log(u1))
    return r, r * mp.cos(2 * mp.pi * u2), r * mp.sin(2 * mp.pi * u2)


def test_box_muller_transform_against_mpmath():
    """The table-driven FP64 transform (plom_bm_pair) against the exact transform: random
    words, the extreme words (u1 -> 2^-33 and u1 -> 1), and every table boundary of both
    words.  The error bound is 1e-15 absolute per unit of radius (a few ulps)."""
    rng = np.random.default_rng(11)
    xs = list(rng.integers(0, 2 ** 32, 3000, dtype=np.uint64))
    ys = list(rng.integers(0, 2 ** 32, 3000, dtype=np.uint64))
    edges = [0, 1, 2, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 2, 2 ** 32 - 1]
    # log table boundaries: u = 2^e (1 + k/128) and the k = 64 switch, at several exponents
    for e in (1, 5, 17, 31):
        for k in (0, 1, 63, 64, 65, 127):
            v = (2 ** e) * (128 + k) // 128
            edges += [v - 1, v, v + 1]
    # angle table boundaries: y = k 2^24 (+-1) and the centre of each entry
    sc = []
    for k in (0, 1, 63, 64, 127, 128, 191, 192, 255):
        sc += [k * 2 ** 24, k * 2 ** 24 + 2 ** 23, (k + 1) * 2 ** 24 - 1]
    xs += edges + [int(rng.integers(0, 2 ** 32)) for _ in sc]
    ys += [int(rng.integers(0, 2 ** 32)) for _ in edges] + sc
    xy = np.array([[int(a) % 2 ** 32, int(b) % 2 ** 32] for a, b in zip(xs, ys)], dtype=np.uint32)
    z = oracle.bm_pairs(xy)
    worst = 0.0
    for (x, y), (z0, z1) in zip(xy, z):
        r, e0, e1 = bm_exact(x, y)
        tol = 1e-15 * max(1.0, float(r))
        worst = max(worst, abs(z0 - float(e0)) / tol, abs(z1 - float(e1)) / tol)
    assert worst <= 1.0, worst


def test_box_muller_near_one_keeps_absolute_accuracy():
    """u1 -> 1 gives a tiny radius; log u1 there must not lose absolute accuracy (the
    table's k >= 64 entries hold log c_k - ln 2 for that reason)."""
    xy = np.array([[2 ** 32 - 1 - d, 2 ** 30] for d in range(0, 4000, 7)], dtype=np.uint32)
    z = oracle.bm_pairs(xy)
    for (x, y), (z0, z1) in zip(xy, z):
        r, e0, e1 = bm_exact(x, y)
        # within 2 ulps of the radius (4.5e-16 relative), also where r ~ 1e-5
        assert abs(z0 - float(e0)) <= 4.5e-16 * float(r) + 1e-22
        assert abs(z1 - float(e1)) <= 4.5e-16 * float(r) + 1e-22


def test_normal_streams_are_the_transform_of_xoshiro_words():
    """oracle.normals(seed, stream, n, j) = the inverse-CDF transform of the xoshiro128**
    words seeded by Philox(ctr = (j, n, 0, 0), key = (seed, stream)), one word each, in
    order."""
    seed, stream, n, j = 9, 4, 17, 300
    w = oracle.philox([j, n, 0, 0], [seed, stream])
    words, _ = oracle.xoshiro(w, 40)
    z = oracle.normals(seed, stream, n, j, 40)
    np.testing.assert_array_equal(z, oracle.icdf(words))


def icdf_exact(w):
    """+-Phi^-1((v + 1/2) 2^-32), v = w >> 1, positive for odd w, in 40-digit arithmetic."""
    u = (mp.mpf(int(w) >> 1) + mp.mpf(1) / 2) / 2 ** 32
    z = mp.sqrt(2) * mp.erfinv(2 * u - 1)
    return -z if int(w) & 1 else z


def test_icdf_transform_against_mpmath():
    """The stream's Gaussian transform (plom_icdf_normal) against the exact inverse normal
    CDF: random words, the extreme words (u -> 2^-33 and u -> 1/2), and the first and last
    word of every polynomial segment (binary octave x 16 sub-segments) with its neighbours.
    Bound: 4e-14 absolute (the fits are 2e-14; |z| <= 6.24)."""
    rng = np.random.default_rng(12)
    ws = [int(x) for x in rng.integers(0, 2 ** 32, 3000, dtype=np.uint64)]
    ws += [0, 1, 2, 3, 4, 5, 2 ** 32 - 1, 2 ** 32 - 2, 2 ** 32 - 3, 2 ** 32 - 4]
    for e in range(-32, -1):                  # segment starts: u = 2^e (1 + s/16), v = u 2^32 - 1/2
        for s_ in range(16):
            v = int((2 ** (e + 32)) * (16 + s_) // 16)
            for d in (-1, 0, 1):
                if 0 <= v + d < 2 ** 31:
                    ws += [2 * (v + d), 2 * (v + d) + 1]
    z = oracle.icdf(np.array(ws, dtype=np.uint64) % 2 ** 32)
    worst = max(abs(zi - float(icdf_exact(w))) for w, zi in zip(ws, z))
    assert worst <= 4e-14, worst


def test_icdf_symmetry_and_range():
    """z(w) = -z(w ^ 1) exactly; the extremes are +-Phi^-1(2^-33) = -+6.2376..."""
    rng = np.random.default_rng(13)
    w = rng.integers(0, 2 ** 31, 20000, dtype=np.uint64) * 2
    np.testing.assert_array_equal(oracle.icdf(w), -oracle.icdf(w + 1))
    e = oracle.icdf(np.array([0, 1], dtype=np.uint32))
    assert e[0] == pytest.approx(float(mp.sqrt(2) * mp.erfinv(2 * mp.mpf(2) ** -33 - 1)), abs=4e-14)
    assert e[1] == -e[0]


def test_icdf_distribution():
    """One million Gaussians of the particle streams: Kolmogorov-Smirnov against N(0, 1)
    and the first four moments within their sampling error."""
    from scipy import stats
    z = np.concatenate([oracle.normals(3, 7, n, j, 1000) for n in range(10) for j in range(100)])
    assert stats.kstest(z, "norm").pvalue > 1e-3
    m = len(z)
    assert abs(z.mean()) < 5 / np.sqrt(m)
    assert abs(z.var() - 1) < 5 * np.sqrt(2 / m)
    assert abs(stats.skew(z)) < 5 * np.sqrt(6 / m)
    assert abs(stats.kurtosis(z)) < 5 * np.sqrt(24 / m)


def test_log_tab_against_mpmath():
    """The psr samplers' table logarithm (plom_log_tab) over the normal doubles: random
    mantissas at exponents -1020..1020, every table boundary (m = 1 + k/128, the k = 64
    switch) and the neighbourhood of 1, where the error must stay absolute (no
    cancellation).  Bound: 2 ulps of log x, or 2.5e-16 absolute below |log x| = 1."""
    rng = np.random.default_rng(5)
    xs = list(np.ldexp(1.0 + rng.random(3000), rng.integers(-1020, 1021, 3000)))
    for e in (-1000, -60, -1, 0, 1, 2, 700):
        for k in (0, 1, 63, 64, 65, 127):
            v = math.ldexp(1.0 + k / 128.0, e)
            xs += [np.nextafter(v, 0.0), v, np.nextafter(v, np.inf)]
    xs += list(1.0 + rng.uniform(-2 ** -7, 2 ** -7, 1000)) + [1.0 - 2 ** -53, 1.0 + 2 ** -52]
    x = np.array(xs, dtype=np.float64)
    got = oracle.log_tab(x)
    for xi, g in zip(x, got):
        ref = mp.log(mp.mpf(float(xi)))
        err = abs(mp.mpf(float(g)) - ref)
        tol = 2.5e-16 if abs(ref) < 1 else 2.0 * math.ulp(float(ref))
        assert err <= tol, (float(xi), float(g), float(ref))
