"""The benchmark workloads: models, parameters, initial states and sizes.

C1  SIR, 1 x 1e6, beta 0.5/d, gamma 0.2/d, rep 0.6, phi 0.1, N = 52 weeks, J = 1000, FP64
C2  SIRS with seasonal forcing (amplitude 0.3, period 365 d) and white noise on beta
    (sigma 0.1), 4 coupled populations of 5e5, N = 520 weeks, J = 1e5, FP64 vs FP32
C3  256 C2-model filters, parameters drawn around the C2 truth, J ~ U{1000..20000},
    N ~ U{104..520}
C4  served stream of SIR / SIRS requests, J log-uniform in [500, 5e4], N ~ U{52..520}
C5  SEIR (waning immunity) with 10 gravity-coupled populations of 1e6, N = 1040 weeks,
    J = 1e6, ancestry kept for --full

Choices the workload names leave open (initial proportions, waning and recovery rates,
coupling matrices, seeds) are fixed here once and used by tests and bench alike.
"""
import numpy as np

DATA_SEED = 20261203      # synthetic series
FILTER_SEED = 1           # random key of the filters

TRUTH = {
    "sir": {"beta": 0.5, "gamma": 0.2, "rep": 0.6, "phi": 0.1},
    "sirs4": {"beta": 0.5, "e": 0.3, "gamma": 0.2, "omega": 1.0 / 365.0, "sto": 0.1, "rep": 0.6, "phi": 0.1},
    "seir10": {"beta": 0.6, "eps": 0.4, "gamma": 0.25, "omega": 1.0 / 730.0, "sto": 0.1, "rep": 0.6, "phi": 0.1},
}

X0_PROPS = {
    "sir": {"S": 0.999, "I": 0.001},
    "sirs4": {"S": [0.45, 0.46, 0.44, 0.47], "I": [0.002, 0.001, 0.003, 0.0015]},
    "seir10": {"S": [0.42 + 0.005 * k for k in range(10)], "E": [0.0005] * 10, "I": [0.001 + 0.0002 * k for k in range(10)]},
}

DATA_N = {"sir": 520, "sirs4": 520, "seir10": 1040}

CONFIGS = {
    "C1": {"model": "sir", "J": 1000, "N": 52, "precision": "fp64"},
    "C2": {"model": "sirs4", "J": 100000, "N": 520, "precision": "fp64"},
    "C2_fp32": {"model": "sirs4", "J": 100000, "N": 520, "precision": "fp32"},
    "C5": {"model": "seir10", "J": 1000000, "N": 1040, "precision": "fp64"},
}


def c3_draws(n_filters=256, seed=3, size_seed=4):
    """Parameter sets around the C2 truth (5 % sd on the log scale) and uneven sizes."""
    rng = np.random.default_rng(seed)
    srng = np.random.default_rng(size_seed)
    out = []
    for i in range(n_filters):
        p = dict(TRUTH["sirs4"])
        for k in ("beta", "gamma", "sto", "rep"):
            p[k] = float(p[k] * np.exp(0.05 * rng.standard_normal()))
        J = int(srng.integers(1000, 20001))
        N = int(srng.integers(104, 521))
        out.append({"params": p, "J": J, "N": N, "stream": 1000 + i})
    return out
