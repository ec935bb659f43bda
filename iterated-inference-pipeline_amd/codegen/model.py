"""Model description -> model semantics (states, reactions, observations, layouts).

Follows the reference's parsing and rewriting rules:
  * states, remainder, fixed / process / observation parameters, white noise
    attachment, observed-variable sort (incidences first), rate rewriting
    (remainder -> (N-S-...), and N -> (S+I+...) when there is no remainder):
    model_builder/Cmodel.py:31-184;
  * populations, cities/ages and time-series naming ``name__stream__inc|prev``:
    model_builder/Context.py:32-81;
  * time-series sort (stream, then name, then observed variable) and obs2ts:
    model_builder/Builder.py:70-169;
  * the ``.settings.json`` written for the C side: model_builder/Builder.py:173-259.

School terms (``terms_forcing(amplitude)`` in a rate, special_functions.c:21-47): the
holiday periods of every population, as fractions of a year, come from the context's
``data.school_terms`` (one flat list ``[begin_0, end_0, begin_1, end_1, ...]`` per
population; the reference's loader reads the same list from the settings,
C/core/build.c:834-868, and its builder leaves writing it to the user, Builder.py:219-220).
They travel in pv, one record of PLOM_SCHOOL_REC values per population after the other
blocks: n_terms, prop_school = 1 - sum(end - begin), then the PLOM_MAX_TERMS (begin, end)
pairs, unused pairs 0.

Extension (absent from the reference grammar, where every state maps to its own
population, model_builder/Ccoder.py:36-37): ``mix(X)`` inside a rate evaluates, for
population ``cac``, sum_k C[cac][k] * X[k] with a coupling matrix C given in the context
as ``"coupling": {"matrix": [[...], ...]}``.  It expresses coupled and gravity-coupled
populations.
"""
import copy
import datetime

from . import expr as E


class ModelError(ValueError):
    pass


MAX_TERMS = 8                    # PLOM_MAX_TERMS: holiday periods per population
SCHOOL_REC = 2 + 2 * MAX_TERMS   # PLOM_SCHOOL_REC: values per population in pv


class PlomModel:
    def __init__(self, context, process, link, name=None):
        self.name = name or process.get("name", "model")
        self.context = copy.deepcopy(context)
        self.process = copy.deepcopy(process)
        self.link = copy.deepcopy(link)

        # ---------------------------------------------------------- Context.py:32-81
        self.cac_id = [p["id"] for p in context.get("population", [])]
        head = [x.split("__") for x in self.cac_id]
        cities = [h[0] for h in head]
        ages = [h[1] for h in head]
        self.cities_id = [cities[i] for i in sorted(cities.index(c) for c in set(cities))]
        self.ages_id = [ages[i] for i in sorted(ages.index(a) for a in set(ages))]
        self.N_C = len(set(cities))
        self.N_AC = len(set(ages))
        self.N_CAC = self.N_C * self.N_AC
        self.ts_id = [x["id"] for x in context.get("time_series", [])]
        self.map_ts_cac = {x["id"]: list(x["population_id"]) for x in context.get("time_series", [])}

        # ---------------------------------------------------------- Cmodel.py:43-76
        self.remainder = None
        self.par_sv = []
        for x in process["state"]:
            tags = x.get("tag", [])
            if isinstance(tags, str):
                tags = [tags]
            if "remainder" in tags:
                self.remainder = x["id"]
            else:
                self.par_sv.append(x["id"])
        meta = context.get("metadata", [])
        self.par_fixed = [x["id"] for x in meta if x["id"] == "N"] + [x["id"] for x in meta if x["id"] != "N"]
        self.par_proc = [x["id"] for x in process["parameter"] if x["id"] not in self.par_fixed]
        self.par_obs = [x["id"] for x in link["observation"][0]["parameter"] if x["id"] not in self.par_fixed]

        self.proc_model = copy.deepcopy(process["model"])
        self.white_noise = []
        for j, x in enumerate(process.get("white_noise", [])):
            wn = {"name": "white_noise__%d" % j, "sd": x["sd"]}
            self.white_noise.append(wn)
            for r in self.proc_model:
                for nr in x["reaction"]:
                    if r["from"] == nr["from"] and r["to"] == nr["to"] and ("rate" not in nr or r["rate"] == nr["rate"]):
                        r["white_noise"] = wn

        self.obs_model = copy.deepcopy(link["observation"][0]["model"])
        if self.obs_model.get("distribution", "discretized_normal") != "discretized_normal":
            raise ModelError("only the discretized_normal observation model exists in the reference (Ccoder.py:528)")

        self.drift_par_proc = [x["parameter"] for x in process.get("diffusion", [])]
        self.vol_par_proc = [x["volatility"] for x in process.get("diffusion", [])]
        self.drift_var = ["drift__par_proc__" + x for x in self.drift_par_proc]

        fixed_obs = []
        for k, v in self.obs_model.items():
            if k not in ("distribution", "comment"):
                fixed_obs += [e for e in E.change_user_input(v) if e in self.par_fixed]
        self.par_fixed_obs = [f for f in self.par_fixed if f in set(fixed_obs)]

        # observed variables, incidences first (stable sort), Cmodel.py:115-147
        observed = copy.deepcopy(link["observed"])
        observed.sort(key=lambda x: 0 if isinstance(x["definition"][0], dict) else 1)
        pm = {}
        for r in self.proc_model:
            pm[(r["from"], r["to"])] = r["rate"]
        self.obs_var, self.obs_var_def, self.obs_ts = [], [], {}
        for x in observed:
            self.obs_var.append(x["id"])
            d = copy.deepcopy(x["definition"])
            for e in d:
                if isinstance(e, dict) and "rate" not in e:
                    e["rate"] = pm[(e["from"], e["to"])]
            self.obs_var_def.append(d)
            self.obs_ts[x["id"]] = list(x.get("time_series_id", []))
        for i, lst in enumerate(self.obs_var_def):
            for k, r in enumerate(lst):
                if isinstance(r, dict):
                    for j, x in enumerate(process.get("white_noise", [])):
                        for nr in x["reaction"]:
                            if r["from"] == nr["from"] and r["to"] == nr["to"] and ("rate" not in nr or r["rate"] == nr["rate"]):
                                r["white_noise"] = {"name": "white_noise__%d" % j, "sd": x["sd"]}

        # rate rewriting, Cmodel.py:150-182
        def resolve_N(tok):
            return "(%s)" % "+".join(self.par_sv) if tok == "N" else tok

        def resolve_rem(tok):
            return "(N-%s)" % "-".join(self.par_sv) if tok == self.remainder else tok

        if not self.remainder:
            for k, v in list(self.obs_model.items()):
                if k != "distribution" and k != "comment":
                    self.obs_model[k] = "".join(map(resolve_N, E.change_user_input(v)))
        for r in self.proc_model:
            if self.remainder:
                if r["from"] == self.remainder:
                    r["rate"] = "(%s)*%s" % (r["rate"], self.remainder)
                r["rate"] = "".join(map(resolve_rem, E.change_user_input(r["rate"])))
            else:
                r["rate"] = "".join(map(resolve_N, E.change_user_input(r["rate"])))
        for lst in self.obs_var_def:
            for d in lst:
                if isinstance(d, dict):
                    if self.remainder:
                        if d["from"] == self.remainder:
                            d["rate"] = "(%s)*%s" % (d["rate"], self.remainder)
                        d["rate"] = "".join(map(resolve_rem, E.change_user_input(d["rate"])))
                    else:
                        d["rate"] = "".join(map(resolve_N, E.change_user_input(d["rate"])))

        # ---------------------------------------------------------- Builder.py:83-169
        self.map_ts_obs = {}
        for x in link["observed"]:
            for ts in x.get("time_series_id", []):
                self.map_ts_obs[ts] = x["id"]
        head = [t.split("__") for t in self.ts_id]
        name_ts = [h[0] for h in head]
        stream = [h[1] for h in head]
        obs_type = [h[2] for h in head]
        tbs = list(zip(self.ts_id, stream, name_ts))
        tbs.sort(key=lambda x: x[1])
        tbs.sort(key=lambda x: x[2])
        tbs.sort(key=lambda x: self.obs_var.index(self.map_ts_obs[x[0]]))
        order = [self.ts_id.index(x[0]) for x in tbs]
        self.ts_id = [self.ts_id[i] for i in order]
        self._name_ts = [name_ts[i] for i in order]
        self._stream = [stream[i] for i in order]
        self._obs_type = [obs_type[i] for i in order]
        self.N_TS = len(self.ts_id)
        self.N_TS_INC = sum(1 for t in self._obs_type if t == "inc")
        self.N_TS_INC_UNIQUE = len(set(n for n, t in zip(self._name_ts, self._obs_type) if t == "inc"))
        self._ts_perm = order  # sorted index -> index in the context's time_series list
        self.obs2ts = self._get_obs2ts()

        self.is_inc = [isinstance(d[0], dict) for d in self.obs_var_def]
        self.N_OBS_INC = sum(self.is_inc)
        self.N_OBS_PREV = len(self.obs_var_def) - self.N_OBS_INC
        n_inc_unique = sum(o["n_ts_unique"] for o, inc in zip(self.obs2ts, self.is_inc) if inc)
        if n_inc_unique != self.N_TS_INC_UNIQUE:
            raise ModelError("incidence time series do not map one-to-one onto observed incidences")

        self.N_PAR_SV = len(self.par_sv)
        # drifting parameters (diffusions, Cmodel.py:84-108): one state slot per population for
        # each, after the states (populations of one group of the parameter share one path)
        self.N_DRIFT = len(self.drift_par_proc)
        self.drift_offset = self.N_PAR_SV * self.N_CAC
        self.inc_offset = self.drift_offset + self.N_DRIFT * self.N_CAC
        self.size_proj = self.inc_offset + self.N_TS_INC_UNIQUE

        coupling = context.get("coupling")
        self.coupling = None
        if coupling is not None:
            m = coupling["matrix"]
            if len(m) != self.N_CAC or any(len(r) != self.N_CAC for r in m):
                raise ModelError("coupling matrix must be N_CAC x N_CAC")
            self.coupling = [[float(v) for v in r] for r in m]

        # rate caches: unique rewritten rates in order of appearance (Ccoder.py:376-383)
        self.rates = []
        for r in self.proc_model:
            if r["rate"] not in self.rates:
                self.rates.append(r["rate"])
        for i, r in enumerate(self.proc_model):
            r["ind_cache"] = self.rates.index(r["rate"])
            r["ind_dem_sto"] = i
        for lst in self.obs_var_def:
            for d in lst:
                if isinstance(d, dict):
                    ids = [i for i, r in enumerate(self.proc_model)
                           if r["from"] == d["from"] and r["to"] == d["to"] and r["rate"] == d["rate"]]
                    d["reactions"] = ids
        self.rate_trees = [E.parse(r) for r in self.rates]
        self.is_school_terms = any(c.fn == "terms_forcing" for t in self.rate_trees for c in E.calls(t))
        self.school_terms = self._school_terms() if self.is_school_terms else None
        self.layout = self._param_layout()

    def check_codegen(self):
        """Raise ModelError for what the code generator does not support."""
        for p, v in zip(self.drift_par_proc, self.vol_par_proc):
            if p not in self.par_proc or v not in self.par_proc:
                raise ModelError("a diffusion needs a process parameter and a process volatility (%s, %s)" % (p, v))
        self._check_symbols()

    # ------------------------------------------------------------------ helpers
    def _get_obs2ts(self):
        out = []
        for o in self.obs_var:
            ind = [i for i, t in enumerate(self.ts_id) if self.map_ts_obs[t] == o]
            names = [self._name_ts[i] for i in ind]
            uniq = []
            for n in names:
                if n not in uniq:
                    uniq.append(n)
            info = {"n_ts_unique": len(uniq), "n_stream": [], "n_cac": [], "cac": []}
            for u in uniq:
                iu = [ind[k] for k, n in enumerate(names) if n == u]
                info["n_stream"].append(len(set(self._stream[i] for i in iu)))
                cacs = self.map_ts_cac[self.ts_id[iu[0]]]
                info["n_cac"].append(len(cacs))
                info["cac"].append([[self.cities_id.index(c.split("__")[0]), self.ages_id.index(c.split("__")[1])]
                                    for c in cacs])
            out.append(info)
        off = 0
        for x in out:
            x["offset"] = off
            off += sum(x["n_stream"])
        return out

    def _check_symbols(self):
        known = set(self.par_sv) | set(self.par_proc) | set(self.par_fixed) | {"t", "M_PI", "ONE_YEAR"}
        for t in self.rate_trees:
            for s in E.symbols(t):
                if s not in known:
                    raise ModelError("unknown symbol %r in rate" % s)
            for c in E.calls(t):
                if c.fn == "terms_forcing" and len(c.args) != 1:
                    raise ModelError("terms_forcing() takes one argument, the amplitude")
                if c.fn == "mix" and self.context.get("coupling") is None:
                    raise ModelError("mix() needs a coupling matrix in the context")

    def _school_terms(self):
        """Per population, the flat list of holiday periods [begin, end, ...] in years
        (context data.school_terms), checked."""
        terms = self.context.get("data", {}).get("school_terms")
        if terms is None:
            raise ModelError("terms_forcing() needs data.school_terms in the context (one list per population)")
        if len(terms) != self.N_CAC:
            raise ModelError("data.school_terms needs one list per population (%d)" % self.N_CAC)
        out = []
        for cac, lst in enumerate(terms):
            v = [float(x) for x in lst]
            if len(v) % 2 or len(v) // 2 > MAX_TERMS:
                raise ModelError("school_terms[%d]: pairs (begin, end), at most %d of them" % (cac, MAX_TERMS))
            for k in range(0, len(v), 2):
                if not (0.0 <= v[k] <= v[k + 1] <= 1.0):
                    raise ModelError("school_terms[%d]: a period [begin, end] within [0, 1] years" % cac)
            out.append(v)
        return out

    def school_records(self):
        """The pv values of the school block: per population n_terms, prop_school (the
        proportion of the year with school in, build.c:860-866), the (begin, end) pairs."""
        out = []
        for v in self.school_terms:
            rec = [0.0] * SCHOOL_REC
            rec[0] = float(len(v) // 2)
            held = 0.0
            for k in range(0, len(v), 2):
                held += v[k + 1] - v[k]
            rec[1] = 1.0 - held
            rec[2:2 + len(v)] = v
            out += rec
        return out

    def _param_layout(self):
        """Offsets of the flat parameter vector pv shared by host and kernels."""
        off = 0
        lay = {"proc": {}, "obs": {}, "fixed": {}, "fixed_obs": {}}
        for p in self.par_proc:
            lay["proc"][p] = off
            off += self.N_CAC
        for p in self.par_obs:
            lay["obs"][p] = off
            off += self.N_TS
        for p in self.par_fixed:
            lay["fixed"][p] = off
            off += self.N_CAC
        for p in self.par_fixed_obs:
            lay["fixed_obs"][p] = off
            off += self.N_TS
        lay["npp"] = len(self.par_proc) * self.N_CAC + len(self.par_obs) * self.N_TS  # the per-particle slots
        lay["fix_off"] = lay["npp"]  # the covariate slots follow them
        lay["n_fix"] = len(self.par_fixed) * self.N_CAC + len(self.par_fixed_obs) * self.N_TS
        lay["coupling"] = off if self.coupling is not None else None
        if self.coupling is not None:
            off += self.N_CAC * self.N_CAC
        lay["school"] = off if self.is_school_terms else None
        if self.is_school_terms:
            off += self.N_CAC * SCHOOL_REC
        lay["size"] = off
        return lay

    def ts_cacs(self):
        """For every sorted time series: list of cac indices it aggregates."""
        out = []
        for t in self.ts_id:
            out.append([self.cities_id.index(c.split("__")[0]) * self.N_AC + self.ages_id.index(c.split("__")[1])
                        for c in self.map_ts_cac[t]])
        return out

    def noise_list(self, dem, env):
        """Noise names drawn per population per substep, in draw order (core_template.c:282-283)."""
        out = []
        if dem:
            out += ["dem_sto__%d" % i for i in range(len(self.proc_model))]
        if env:
            out += [w["name"] for w in self.white_noise]
        return out

    def settings(self, data=None, dates=None, date_0=None, times=None, par_fixed_values=None):
        """The .settings.json document (schema of Builder.make_settings_json, Builder.py:173-259)."""
        all_order = self.par_sv + self.par_proc + self.par_obs
        s = {
            "POP_SIZE_EQ_SUM_SV": self.remainder is None,
            "data": {
                "obs2ts": self.obs2ts,
                "data": data if data is not None else [],
                "par_fixed_values": par_fixed_values if par_fixed_values is not None else {},
                "dates": dates if dates is not None else [],
                "times": times if times is not None else [0],
                **({"school_terms": self.school_terms} if self.is_school_terms else {}),
            },
            "drift": {"ind_par_Xdrift_applied": [all_order.index(x) for x in self.drift_par_proc],
                      "ind_volatility_Xdrift": [all_order.index(x) for x in self.vol_par_proc]},
            "ind_noise_sd": sorted(set(all_order.index(w["sd"]) for w in self.white_noise if w["sd"] in all_order)),
            "remainder": self.remainder,
            "date_0": str(date_0) if date_0 is not None else None,
            "cst": {
                "N_C": self.N_C, "N_AC": self.N_AC,
                "N_PAR_PROC": len(self.par_proc), "N_PAR_OBS": len(self.par_obs),
                "N_PAR_SV": len(self.par_sv), "N_PAR_FIXED": len(self.par_fixed),
                "N_TS": self.N_TS, "N_TS_INC": self.N_TS_INC, "N_TS_INC_UNIQUE": self.N_TS_INC_UNIQUE,
                "N_DATA": len(data) if data is not None else 0,
                "N_OBS_ALL": len(self.obs_var_def), "N_OBS_INC": self.N_OBS_INC, "N_OBS_PREV": self.N_OBS_PREV,
                "N_DRIFT": self.N_DRIFT, "IS_SCHOOL_TERMS": 1 if self.is_school_terms else 0,
            },
            "orders": {
                "par_sv": self.par_sv, "par_proc": self.par_proc, "par_fixed": self.par_fixed,
                "par_obs": self.par_obs, "ts_id": self.ts_id, "cac_id": self.cac_id, "drift_var": self.drift_var,
            },
            # extension consumed by this build's C host: the model the kernels were generated for
            "model": self.name,
        }
        return s


def context_data(model, context):
    """Dates, times (days since date_0) and the data matrix in sorted time-series order
    (Context.py:62-81, Builder.py:102-107, 228-229)."""
    src = context["data"]["source"]
    ts_ctx = [x["id"] for x in context["time_series"]]
    t0 = src[ts_ctx[0]]["t0"]
    date_0 = datetime.datetime.strptime(t0, "%Y-%m-%d").date()
    dates = [x[0] for x in src[ts_ctx[0]]["value"]]
    data = []
    for i in range(len(dates)):
        data.append([src[ts]["value"][i][1] for ts in model.ts_id])
    times = [0] + [(datetime.datetime.strptime(d, "%Y-%m-%d").date() - date_0).days for d in dates]
    return dates, times, data, date_0
