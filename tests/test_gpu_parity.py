"""GPU filter vs the CPU oracle (canonical arithmetic), through the C ABI.

Bar: bit-exact log-weights, ancestors, per-interval log-likelihoods, ESS and total
log-likelihood in FP64 (the floor the workload states is 1e-9 relative; the tests ask
for equality).
"""
import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs, smc
from iip_amd.smc import Model

pytestmark = pytest.mark.gpu


def setup(name, N, params=None, props=None):
    m = Model.get(name)
    times, data = m.dataset(N)
    pv = m.pack(params or configs.TRUTH[name])
    x0 = m.x0(props or configs.X0_PROPS[name])
    return m, times, data, pv, x0


def compare(ctx, name, J, N, data_edit=None, noises_off=0, flags=0, seed=1, stream=0, like_min=1e-17, params=None):
    m, times, data, pv, x0 = setup(name, N, params)
    if data_edit:
        data = data_edit(data.copy())
    print("case", name, "J", J, "N", N, "noises_off", noises_off, "flags", flags, flush=True)
    g = smc.run(ctx, m, J, times, data, pv, x0, ancestry=True, final_state=True, noises_off=noises_off,
                flags=flags, seed=seed, stream=stream, like_min=like_min)
    o = oracle.smc(m.id, J, times, data, pv, x0, noises_off=noises_off, flags=flags, seed=seed, stream=stream,
                   like_min=like_min, ancestry=True, final_state=True, logw=True, nthreads=8)
    assert g.n_all_fail == o["n_all_fail"]
    np.testing.assert_array_equal(g.ancestry, o["ancestry"])
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    np.testing.assert_array_equal(g.ess_n, o["ess_n"])
    assert g.loglik == o["loglik"]
    np.testing.assert_array_equal(g.final_state, o["final_state"])
    return g, o


def test_math_bitwise(ctx):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-40, 40, 20000), rng.uniform(-745, 709, 20000),
                        np.exp(rng.uniform(-700, 700, 20000)), [0.0, 1.0, -1.0, 0.5, 1e-310, 2.0, 38.0, -38.0]])
    # column 6 (sqrt of the Box-Muller radius) is defined on [2^-767, 2^1023] only
    np.testing.assert_array_equal(smc.selftest_math(ctx, x)[:, :6], oracle.math(x)[:, :6])


def test_sqrt_mid_bitwise(ctx):
    """The device sqrt of the Box-Muller radius (v_rsq_f64 + Newton core without range
    scaling) equals IEEE sqrt on 1.5e6 values spread over its argument range and beyond it
    down to 2^-767, where the core is still the compiler's own sequence."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(2.3e-10, 46.1, 1000000), np.exp2(rng.uniform(-767, 1023, 500000)),
                        [2.0 ** -767, 2.3e-10, 46.1, 1.0, 4.0, 2.0]])
    got = smc.selftest_math(ctx, x)
    np.testing.assert_array_equal(got[:, 5], np.sqrt(x))
    np.testing.assert_array_equal(got[:, 6], np.sqrt(x))


def test_sqrt_rate_bitwise(ctx):
    """The device's demographic-noise square root (Newton core and one integer compare)
    equals the oracle's definition on its whole domain: normal and subnormal arguments on
    both sides of 2^-767, zeros, negative numbers, infinities and NaN."""
    rng = np.random.default_rng(9)
    x = np.concatenate([np.exp2(rng.uniform(-1074, 1023, 1000000)), rng.uniform(0, 1e6, 200000),
                        -np.exp2(rng.uniform(-1074, 1023, 1000)),
                        [0.0, -0.0, 5e-324, 2.0 ** -767, np.nextafter(2.0 ** -767, 0.0), np.nextafter(2.0 ** -767, 1.0),
                         np.finfo(float).max, np.inf, -np.inf, np.nan]])
    np.testing.assert_array_equal(smc.selftest_math(ctx, x)[:, 7], oracle.math(x)[:, 7])


def test_normals_bitwise(ctx):
    for (seed, stream, n0, j) in [(1, 0, 0, 0), (7, 3, 11, 12345), (0xffffffff, 0xffffffff, 1039, 999999)]:
        np.testing.assert_array_equal(smc.selftest_normals(ctx, seed, stream, n0, j, 257),
                                      oracle.normals(seed, stream, n0, j, 257))


def test_fp32_normals_are_the_fp64_normals_rounded(ctx):
    """The FP32 stream draws the FP64 stream's Gaussians: the same words through the same
    inverse-CDF segments in FP32, from the top 24 bits of each word (|u32 - u64| <= 2^-25),
    so |z32 - z64| <= 2^-25 / phi(z) + FP32 rounding."""
    from scipy.stats import norm
    for (seed, stream, n0, j) in [(1, 0, 0, 0), (7, 3, 11, 12345), (0xffffffff, 0xffffffff, 1039, 999999)]:
        z64 = oracle.normals(seed, stream, n0, j, 20000)
        z32 = smc.selftest_normals32(ctx, seed, stream, n0, j, 20000)
        bound = 2.0 ** -25 / norm.pdf(np.maximum(np.abs(z32), np.abs(z64))) * 1.01 + 1e-6
        assert np.all(np.abs(z32 - z64) <= bound)
        assert np.all((z32 >= 0) == (z64 >= 0))


def test_sir_j1000_n52_parity(ctx):
    """The CPU-runnable SIR workload exactly: J = 1000, N = 52, FP64."""
    compare(ctx, "sir", 1000, 52)


@pytest.mark.parametrize("J", [1, 7, 256, 257, 1000, 4099])
def test_sir_ragged_J(ctx, J):
    compare(ctx, "sir", J, 20)


def test_sirs4_parity(ctx):
    compare(ctx, "sirs4", 3000, 40)


def test_seir10_parity(ctx):
    compare(ctx, "seir10", 1500, 20)


def test_many_tiles(ctx):
    """More than 256 tiles: the tile-level tree runs over several chunks."""
    compare(ctx, "sir", 70000, 6)


@pytest.mark.parametrize("noises_off", [1, 2, 3])
def test_noise_variants(ctx, noises_off):
    compare(ctx, "sirs4", 600, 12, noises_off=noises_off)


def test_missing_data(ctx):
    """NaN series and all-NaN rows (no weighting, identity ancestry)."""
    def edit(d):
        d[2, :] = np.nan
        d[5, 1] = np.nan
        d[6, :2] = np.nan
        return d
    compare(ctx, "sirs4", 900, 10, data_edit=edit)


def test_all_fail_branch(ctx):
    """Data far from the model with a huge LIKE_MIN: every particle floored at some steps."""
    def edit(d):
        d[3, :] = 1e9
        return d
    g, o = compare(ctx, "sir", 500, 8, data_edit=edit, like_min=1e-3)
    assert g.n_all_fail >= 1
    assert np.all(g.ancestry[3] == np.arange(500))


def test_no_filter(ctx):
    g, o = compare(ctx, "sir", 300, 10, flags=smc.FLAG_NO_FILTER)
    assert g.loglik == 0.0


def test_streams_and_seeds(ctx):
    a, _ = compare(ctx, "sir", 700, 10, seed=5, stream=9)
    b, _ = compare(ctx, "sir", 700, 10, seed=5, stream=10)
    assert a.loglik != b.loglik


@pytest.mark.parametrize("flags", [0, smc.FLAG_PSR])
def test_long_intervals_evaluate_the_forcing_inline(ctx, flags):
    """Intervals of more than 512 substeps do not fit the block's time-function table: the
    step evaluates the seasonal forcing itself.  Long and weekly intervals mixed, FP64 equal
    to the oracle (the table and the inline evaluation give the same values)."""
    m = Model.get("sirs4")
    _, data = m.dataset(8)
    times = np.array([0.0, 7.0, 14.0, 150.0, 157.0, 300.0, 307.0, 314.0, 321.0])
    pv, x0 = m.pack(configs.TRUTH["sirs4"]), m.x0(configs.X0_PROPS["sirs4"])
    if flags & smc.FLAG_PSR:
        x0 = np.round(x0)
    print("long intervals", flags, flush=True)
    g = smc.run(ctx, m, 1000, times, data, pv, x0, flags=flags, seed=3, stream=1)
    o = oracle.smc(m.id, 1000, times, data, pv, x0, seed=3, stream=1, nthreads=8, psr=bool(flags & smc.FLAG_PSR))
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    assert g.loglik == o["loglik"]
