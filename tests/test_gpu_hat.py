"""hat (weighted means and 95 % bounds per interval) and prediction residuals on the GPU
against the CPU oracle, which is pinned against numpy in tests/test_oracle.py.  FP64:
bit-exact (the bounds come from exact integer weight sums, the means and residuals from
fixed tree shapes)."""
import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs, smc
from iip_amd.smc import Model

pytestmark = pytest.mark.gpu


def setup(name, N):
    m = Model.get(name)
    times, data = m.dataset(N)
    return m, times, data.copy(), m.pack(configs.TRUTH[name]), m.x0(configs.X0_PROPS[name])


def gpu_hat(ctx, m, J, times, data, pv, x0, precision="fp64", flags=0, seed=3, stream=4):
    f = smc.Filter(ctx, m, J, times, data, pv, x0, precision=precision, flags=flags | smc.FLAG_HAT, seed=seed,
                   stream=stream)
    f.run()
    r = f.read(hat=True)
    f.close()
    return r


@pytest.mark.parametrize("name,J,N,missing,flags", [("sir", 1000, 52, None, 0), ("sirs4", 2500, 30, 7, 0),
                                                    ("seir10", 600, 15, None, 0), ("sir", 300, 20, 5, smc.FLAG_NO_FILTER),
                                                    ("sirs4", 1, 6, None, 0)])
def test_hat_equals_oracle(ctx, name, J, N, missing, flags):
    m, times, data, pv, x0 = setup(name, N)
    if missing is not None:
        data[missing, :] = np.nan
    print("hat", name, J, N, missing, flags, flush=True)
    g = gpu_hat(ctx, m, J, times, data, pv, x0, flags=flags)
    o = oracle.smc(m.id, J, times, data, pv, x0, flags=flags, seed=3, stream=4, hat=True, nthreads=8)
    np.testing.assert_array_equal(g.hat, o["hat"])
    np.testing.assert_array_equal(g.pred_res, o["pred_res"])
    assert g.loglik == o["loglik"]


@pytest.mark.parametrize("lanes", [1, 2])
def test_hat_batched_equals_alone(ctx, lanes):
    cases = [("sir", 700, 25, 1, "fp64"), ("sirs4", 1300, 18, 2, "fp64"), ("sirs4", 900, 22, 3, "fp32")]
    fs, alone = [], []
    for name, J, N, stream, prec in cases:
        m, times, data, pv, x0 = setup(name, N)
        fs.append(smc.Filter(ctx, m, J, times, data, pv, x0, precision=prec, flags=smc.FLAG_HAT, seed=9,
                             stream=stream))
        alone.append(gpu_hat(ctx, m, J, times, data, pv, x0, precision=prec, seed=9, stream=stream))
    b = smc.Batch(ctx, fs)
    b.configure(lanes)
    b.launch(use_graph=True)
    for f, a in zip(fs, alone):
        r = f.read(hat=True)
        np.testing.assert_array_equal(r.hat, a.hat)
        np.testing.assert_array_equal(r.pred_res, a.pred_res)
    b.close()
    for f in fs:
        f.close()
