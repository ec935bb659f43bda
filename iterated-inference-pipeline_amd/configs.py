"""The benchmark workloads: models, parameters, initial states and sizes.

sir_j1000_n52        SIR, 1 x 1e6, beta 0.5/d, gamma 0.2/d, rep 0.6, phi 0.1, N = 52 weeks,
                     J = 1000, FP64 (the CPU-runnable case)
sirs4_j100000_n520   SIRS with seasonal forcing (amplitude 0.3, period 365 d) and white noise
                     on beta (sigma 0.1), 4 coupled populations of 5e5, N = 520 weeks,
                     J = 1e5, FP64 and FP32 propagation
pmcmc_batch          256 SIRS-model filters (pMCMC likelihood evaluations), parameters drawn
                     around the SIRS truth, J ~ U{1000..20000}, N ~ U{104..520}
served_mix           Poisson stream (200 requests/s) of SIR / SIRS requests, J log-uniform in
                     [500, 5e4], N ~ U{52..520}
seir10_j1000000_n1040  SEIR (waning immunity) with 10 gravity-coupled populations of 1e6,
                     N = 1040 weeks, J = 1e6, ancestry kept for --full
noise                the SIR model with births, deaths and white noise of the reference's
                     example (model_builder/example/noise), 2 cities, 4 series (three
                     incidences, one prevalence), with time-varying covariates: population
                     sizes, birth and death rates (per week) and proportions under
                     surveillance, N = 104 weeks
drift                the reference's example model_builder/example/drift: the noise model without
                     white noise and with a diffusing r0 (volatility vol), same context

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
    "noise": {"r0": 1.8, "v": 0.1, "sto": 0.1, "rep": 0.6, "phi": 0.1},
    "drift": {"r0": 1.8, "v": 0.1, "vol": 0.05, "rep": 0.6, "phi": 0.1},
}

X0_PROPS = {
    "sir": {"S": 0.999, "I": 0.001},
    "sirs4": {"S": [0.45, 0.46, 0.44, 0.47], "I": [0.002, 0.001, 0.003, 0.0015]},
    "seir10": {"S": [0.42 + 0.005 * k for k in range(10)], "E": [0.0005] * 10, "I": [0.001 + 0.0002 * k for k in range(10)]},
    "noise": {"S": [0.7, 0.65], "I": [0.001, 0.0005]},
    "drift": {"S": [0.7, 0.65], "I": [0.001, 0.0005]},
}

DATA_N = {"sir": 520, "sirs4": 520, "seir10": 1040, "noise": 104, "drift": 104}

WORKLOADS = {
    "sir_j1000_n52": {"model": "sir", "J": 1000, "N": 52, "precision": "fp64"},
    "sirs4_j100000_n520": {"model": "sirs4", "J": 100000, "N": 520, "precision": "fp64"},
    "sirs4_j100000_n520_fp32": {"model": "sirs4", "J": 100000, "N": 520, "precision": "fp32"},
    "seir10_j1000000_n1040": {"model": "seir10", "J": 1000000, "N": 1040, "precision": "fp64"},
}

def drift_setup(model, x0, params, groups=None):
    """The drift map of a model with diffusions and its initial state: every drifting
    parameter on the log scale (as the example's theta), one group per population unless
    groups [n_drift][n_cac] says otherwise; the volatility of group g is the pv slot of the
    first population of that group; the drift states start at log of the parameter's
    natural value."""
    import math
    m = model.desc
    n = m.N_DRIFT * m.N_CAC
    g = np.asarray(groups if groups is not None else [list(range(m.N_CAC))] * m.N_DRIFT, dtype=np.int32).ravel()
    vols = np.full(n, -1, dtype=np.int32)
    for i, v in enumerate(m.vol_par_proc):
        for gi in range(m.N_CAC):
            cacs = [c for c in range(m.N_CAC) if g[i * m.N_CAC + c] == gi]
            if cacs:
                vols[i * m.N_CAC + gi] = m.layout["proc"][v] + cacs[0]
    drift = {"group": g, "tr": np.ones(n, dtype=np.int32), "scale": np.zeros(n, dtype=np.int32),
             "duration": np.zeros(n, dtype=np.int32), "min": np.zeros(n), "max": np.zeros(n),
             "multiplier": np.ones(n), "vol_slot": vols}
    x = np.array(x0, dtype=np.float64)
    for i, p in enumerate(m.drift_par_proc):
        for c in range(m.N_CAC):
            x[m.drift_offset + i * m.N_CAC + c] = math.log(params[p])
    return x, drift


# parameters a proposal moves (multiplicatively, on the log scale)
PROPOSED = {"sir": ("beta", "gamma", "rep"), "sirs4": ("beta", "gamma", "sto", "rep")}


def proposal(model, rng, sd=0.05):
    """A parameter set around the truth of `model`: sd on the log scale."""
    p = dict(TRUTH[model])
    for k in PROPOSED[model]:
        p[k] = float(p[k] * np.exp(sd * rng.standard_normal()))
    return p


def pmcmc_batch(n_filters=256, seed=3, size_seed=4):
    """The batched pMCMC workload: parameter sets around the SIRS truth (5 % sd on the log
    scale) and uneven sizes J ~ U{1000..20000}, N ~ U{104..520}."""
    rng = np.random.default_rng(seed)
    srng = np.random.default_rng(size_seed)
    out = []
    for i in range(n_filters):
        p = proposal("sirs4", rng)
        J = int(srng.integers(1000, 20001))
        N = int(srng.integers(104, 521))
        out.append({"model": "sirs4", "params": p, "J": J, "N": N, "stream": 1000 + i})
    return out


def served_mix(rate=200.0, duration_s=60.0, seed=5):
    """The served workload: Poisson arrivals at `rate` per second over `duration_s`, half SIR
    and half SIRS requests, J log-uniform in [500, 50000], N ~ U{52..520}, parameters
    around each model's truth.  Returns requests sorted by arrival time (seconds)."""
    rng = np.random.default_rng(seed)
    t, out, i = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration_s:
            return out
        model = "sir" if rng.random() < 0.5 else "sirs4"
        J = int(round(float(np.exp(rng.uniform(np.log(500.0), np.log(50000.0))))))
        N = int(rng.integers(52, 521))
        out.append({"t": t, "model": model, "params": proposal(model, rng), "J": J, "N": N, "stream": 100000 + i})
        i += 1
