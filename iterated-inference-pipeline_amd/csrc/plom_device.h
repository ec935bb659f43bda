/*
 * plom_device.h -- the model-dependent device code of the SMC filter (gfx950): the device
 * filter descriptor, the per-particle kernels templated on a generated model (propagate,
 * lineage, hat tiles, prediction residuals) and their host launchers.
 *
 * Every model is compiled in a translation unit of its own (plom_model.hip), which
 * includes this header and one generated model header and exports one model_entry; the
 * library (plom_smc.hip) holds the model-independent kernels (weights, scan, advance,
 * final resampling, hat selection) and the host side, and reaches a model's kernels only
 * through its model_entry.  A model outside the library is compiled the same way into a
 * shared library of its own and registered at run time (plom_model_load).
 */
#ifndef PLOM_DEVICE_H
#define PLOM_DEVICE_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "plom_smc.h"
#include "plom_math.h"
#include "plom_cov.h"
#include "plom_ode.h"
#include "plom_rng.h"
#include "plom_mif.h"
#include "plom_psr.h"

typedef unsigned __int128 plom_u128;

#define TILE 256
#define PLOM_TF_CAP 512   /* LDS entries of the per-interval time-function table */
#define MAX_TILES (PLOM_MAX_J / TILE)

/* ------------------------------------------------------------------ device filter */
struct plom_dfilter {
    int32_t J, Jp, T, N, S, n_ts, flags, step;
    uint32_t seed, stream, prec, pad;
    double dt, sqdt, like_min, all_fail_llike;
    double eps_abs, eps_rel;   /* PLOM_FLAG_ODE: error control of the RKF45 integration */
    void *X[2];
    const void *x0;
    const double *pv;
    const float *pvf;
    const double *cov;  /* covariate table [N kmax + N + 1][cov_cs] (plom_cov.h), or the covariate
                           slots of pv with cov_cs = 0 (constant covariates) */
    int32_t cov_cs, cov_kmax;
    plom_drift_view dv; /* drifting parameters (used when drift_on) */
    int32_t drift_on, pad3;
    const double *times, *data;
    const int32_t *nonan;
    double *logw, *tmax, *tsum, *tsum2;
    uint64_t *cum, *ttot, *tincl;
    uint32_t *cnt;      /* [0] propagate arrivals, [1] scan arrivals */
    double *scal;       /* m, S, S2, loglik */
    int32_t *n_all_fail;
    double *llike_n, *ess_n;
    int32_t *rflag;     /* rflag[n] = 1 when interval n ended with a resampling */
    int32_t *ancestry;  /* [N][J] or null */
    int32_t *lin_slots; /* [N] or null: the sampled lineage's particle index per interval */
    double *lin_state;  /* [(N+1)][S] or null: its states */
    double *lin_obs;    /* [(N+1)][n_ts][2] or null: its observation mean and one drawn observation */
    void *lin_snap;     /* [lin_nseg - 1][S][Jp] or null: the propagated cloud after intervals K-1,
                           2K-1, ... (K = lin_K), where the lineage replay's segments start */
    int32_t lin_K, lin_nseg;
    int32_t n_hat, pad2;
    uint64_t *hkey;     /* [n_hat][Jp] or null: per tile, the sorted keys of the hat quantities */
    uint64_t *hpref;    /* [n_hat][Jp]: inclusive prefix of their weights in that order */
    double *hpart;      /* [n_hat][T]: per tile, tree sum of value x weight */
    double *hat;        /* [(N+1)][n_hat][3]: low95, mean, high95 */
    double *pres;       /* [N][n_ts][2]: prediction mean and residual */
    double *prw;        /* [T] per tile: the resampled copies (offspring) of its particles */
    double *prp;        /* [4][n_ts][T] per tile and series: sum o mu, sum o var, mean mu, sum o (mu - mean)^2 */
    /* PLOM_FLAG_MIF: per-particle parameters */
    int32_t n_theta, n_pp;
    const int32_t *pm_src;          /* [n_pp] component of each per-particle pv slot, -1 = pv */
    const int32_t *pm_tr, *pm_sc, *pm_dur;      /* [n_theta] */
    const double *pm_min, *pm_max, *pm_mult, *pm_sd;   /* [n_theta] */
    const double *mut_fac;          /* [N] random-walk factor after interval n (with data) */
    const double *pm_chol;          /* [n_theta][n_theta] lower factor of the walk's covariance, or null */
    const int32_t *pm_prior, *pm_puntil;   /* [n_theta] prior patch of the weights, or null */
    const double *pm_ppow;                 /* [n_theta] */
    double *th[2];                  /* [n_theta][Jp] theta of the particles, ping-pong like X */
    double *mw;                     /* [Jp] the weights exp(l - m) of the last resampled interval
                                       (k_scan), read by the theta moments */
    double *mpw, *mps, *mpm, *mpq;  /* tile partials: [T] sum w, [n_theta][T] sum w theta, mean, M2 */
    double *mmean, *mvar;           /* [(N+1)][n_theta] weighted mean and variance of theta */
    void *allX;                     /* PLOM_FLAG_ALL_STATES: [(N+1)][S][Jp] every propagated cloud */
    /* the ode implementation with time-varying covariates: their splines, evaluated at the
     * integrator's times (spl_on = 0: the covariates are constants) */
    plom_spline_set spl;
    int32_t spl_on, pad_spl;
};

/* ------------------------------------------------------------------ covariate rows */
/* the rows of interval n's substeps (the same in every lane: moved to scalar registers so
 * that the generated code reads the values with scalar loads) */
__device__ inline const double *cov_steps(const plom_dfilter *f, int n)
{
    const uint64_t u = (uint64_t)(f->cov + (size_t)n * (size_t)f->cov_kmax * (size_t)f->cov_cs);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return (const double *)(((uint64_t)hi << 32) | lo);
}

/* the row of observation time times[r], 0 <= r <= N */
__device__ inline const double *cov_at(const plom_dfilter *f, int r)
{
    return f->cov + ((size_t)f->N * (size_t)f->cov_kmax + (size_t)r) * (size_t)f->cov_cs;
}

/* ------------------------------------------------------------------ block helpers */
/* recursive contiguous pairwise sum of the 256 values of a block (every lane gets it) */
__device__ inline double tree256(double v, double *lds)
{
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) v = v + __shfl_xor(v, h, 64);
    const int w = (threadIdx.x >> 6) & 3;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double r = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    __syncthreads();
    return r;
}

__device__ inline double max256(double v, double *lds)
{
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) v = fmax(v, __shfl_xor(v, h, 64));
    const int w = (threadIdx.x >> 6) & 3;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double r = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
    __syncthreads();
    return r;
}

/* inclusive scan of 256 uint64 values (exact: any order gives the same integers) */
__device__ inline uint64_t scan256(uint64_t v, uint64_t *lds)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) {
        uint64_t o = __shfl_up(v, h, 64);
        if (lane >= h) v += o;
    }
    if (lane == 63) lds[w] = v;
    __syncthreads();
    uint64_t off = 0;
    for (int k = 0; k < w; k++) off += lds[k];
    __syncthreads();
    return v + off;
}

/* first particle i with C_i * J2 >= Rk in the fixed-point cumulative weights of the
 * interval whose scan is in f->cum / f->tincl */
__device__ inline int select_first(const plom_dfilter *f, plom_u128 Rk, uint64_t J2)
{
    const int T = f->T;
    const uint64_t *tincl = f->tincl;
    int lo = 0, hi = T - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if ((plom_u128)tincl[mid] * J2 >= Rk) hi = mid; else lo = mid + 1;
    }
    const uint64_t P = lo ? tincl[lo - 1] : 0ull;
    const uint64_t *c = f->cum + (size_t)lo * TILE;
    int a = 0, b = TILE - 1;
    while (a < b) {
        int mid = (a + b) >> 1;
        if ((plom_u128)(P + c[mid]) * J2 >= Rk) b = mid; else a = mid + 1;
    }
    return lo * TILE + a;
}

/* systematic resampling of interval n: ancestor of output slot k, the first i with
 * C_i J 2^33 >= (k 2^33 + 2x + 1) Q */
__device__ inline int select_ancestor(const plom_dfilter *f, int k, int n)
{
    const uint64_t Q = f->tincl[f->T - 1];
    const uint64_t unum = plom_resample_num(f->seed, f->stream, (uint32_t)n);
    const plom_u128 Rk = (plom_u128)(((uint64_t)k << 33) + unum) * (plom_u128)Q;
    return select_first(f, Rk, (uint64_t)f->J << 33);
}


/* ------------------------------------------------------------------ kernels */
/* PLOM_FLAG_MIF: theta of particle j from its ancestor a, moved by the random walk when
 * interval n-1 had data (resample_and_mut_theta_mif, C/mif/methods.c:261-367: every
 * component with a jump size moves by mut_fac[n-1] * sd * N(0,1)); written to the output
 * table; then the per-particle slots of pv from it (back_transform_theta2par).  The slots
 * take each component's natural value as the component is moved, from the register that
 * holds it (the slot map pm_src is uniform, so the match is a scalar branch), instead of
 * reading the output table back. */
template <int NPP>
__device__ inline void mif_slots(const plom_dfilter *f, const int *src, int o, double v, double *lp)
{
    bool used = false;
#pragma unroll
    for (int s = 0; s < NPP; s++) used = used || src[s] == o;
    if (!used) return;
    const double nat = plom_mif_natural(f->pm_tr[o], f->pm_sc[o], f->pm_dur[o], v, f->pm_min[o], f->pm_max[o], f->pm_mult[o]);
#pragma unroll
    for (int s = 0; s < NPP; s++)
        if (src[s] == o) lp[s] = nat;
}

template <int NPP>
__device__ inline void mif_particle(const plom_dfilter *f, int n, int j, int a, const double *tab, double *lp)
{
    const int Jp = f->Jp, nth = f->n_theta;
    const double *Tin = f->th[n & 1];
    double *Tout = f->th[(n + 1) & 1];
    const double fac = (n > 0 && f->nonan[n - 1] > 0) ? f->mut_fac[n - 1] : 0.0;
    int src[NPP];
#pragma unroll
    for (int s = 0; s < NPP; s++) {
        src[s] = f->pm_src[s];
        lp[s] = f->pv[s];          /* slots without a component keep pv */
    }
    if (f->pm_chol) {
        /* correlated walk: theta_o += fac sum_{q<=o} L[o][q] z_q, z from the same counters */
        double z[PLOM_MIF_MVN_MAX];
        for (int o = 0; o < nth; o += 2) {
            double z0 = 0.0, z1 = 0.0;
            if (fac != 0.0) plom_mif_normals(f->seed, f->stream, (uint32_t)n, (uint32_t)j, (uint32_t)(o >> 1), tab, &z0, &z1);
            z[o] = z0;
            if (o + 1 < nth) z[o + 1] = z1;
        }
        for (int o = 0; o < nth; o++) {
            double v = Tin[(size_t)o * Jp + a];
            if (fac != 0.0) {
                double acc = 0.0;
                for (int q = 0; q <= o; q++) acc = acc + f->pm_chol[(size_t)o * nth + q] * z[q];
                v = v + fac * acc;
            }
            Tout[(size_t)o * Jp + j] = v;
            mif_slots<NPP>(f, src, o, v, lp);
        }
    } else
    for (int o = 0; o < nth; o += 2) {
        const double sd0 = fac * f->pm_sd[o];
        const double sd1 = (o + 1 < nth) ? fac * f->pm_sd[o + 1] : 0.0;
        double z0 = 0.0, z1 = 0.0;
        if (sd0 != 0.0 || sd1 != 0.0) plom_mif_normals(f->seed, f->stream, (uint32_t)n, (uint32_t)j, (uint32_t)(o >> 1), tab, &z0, &z1);
        double v = Tin[(size_t)o * Jp + a];
        if (sd0 != 0.0) v = v + sd0 * z0;
        Tout[(size_t)o * Jp + j] = v;
        mif_slots<NPP>(f, src, o, v, lp);
        if (o + 1 < nth) {
            double w = Tin[(size_t)(o + 1) * Jp + a];
            if (sd1 != 0.0) w = w + sd1 * z1;
            Tout[(size_t)(o + 1) * Jp + j] = w;
            mif_slots<NPP>(f, src, o + 1, w, lp);
        }
    }
}

/* One observation interval for TPB tiles of 256 particles per block (blocks of 256 TPB
 * threads): sub-block q = threadIdx.x / 256 takes entry blockIdx.x TPB + q of the tile map
 * (or that tile of the single filter without a map) and has its own time-function table
 * and tile statistics; with TPB > 1 every thread passes the same barriers, whatever its
 * sub-block does.  The tile of the statistics stays 256 particles for every TPB: it fixes
 * the summation tree of the weights (the canonical arithmetic). */
template <class M, typename R, int V, bool MIF, int TPB>
__global__ __launch_bounds__(TILE * TPB) void k_propagate(plom_dfilter *filters, const int2 *tiles, int ntiles)
{
    __shared__ double lds[4 * TPB];
    const int q = TPB > 1 ? (int)(threadIdx.x >> 8) : 0;
    const int tid = TPB > 1 ? (int)(threadIdx.x & (TILE - 1)) : (int)threadIdx.x;
    const int e = (int)blockIdx.x * TPB + q;
    plom_dfilter *f = nullptr;
    int tile = e, n = 0;
    bool act = e < ntiles;
    if (act) {
        int fi = 0;
        if (tiles) { const int2 tt = tiles[e]; fi = tt.x; tile = tt.y; }
        f = &filters[fi];
        n = f->step;
        act = n < f->N;
    }
    if constexpr (TPB == 1) {
        if (!act) return;      /* the whole block */
    }
    /* in LDS: the coefficient table of the stream's Gaussians (not for the ode
     * implementation, which draws none; FP64 or its float copy), with MIF in FP64 the
     * Box-Muller tables of the random walk, and for psr the log and log k! tables of its
     * samplers */
    constexpr bool PSR = V == 4 || V == 5;
    __shared__ double s_icdf[PLOM_ICDF_TAB_SIZE];
    __shared__ float s_icdf32[PLOM_ICDF_TAB_SIZE];
    __shared__ double s_bm[PLOM_BM_TAB_SIZE];
    __shared__ double s_lf[PSR ? 30 : 1];
    if constexpr (sizeof(R) == 8 && V != 6) {
        for (int i = threadIdx.x; i < PLOM_ICDF_TAB_SIZE; i += TILE * TPB) s_icdf[i] = plom_icdf_tab[i];
    }
    if constexpr (sizeof(R) == 4) {
        for (int i = threadIdx.x; i < PLOM_ICDF_TAB_SIZE; i += TILE * TPB) s_icdf32[i] = plom_icdf_tab32[i];
    }
    if constexpr (sizeof(R) == 8 && (MIF || PSR)) {
        for (int i = threadIdx.x; i < PLOM_BM_TAB_SIZE; i += TILE * TPB) s_bm[i] = plom_bm_tab[i];
    }
    if constexpr (PSR) {
        for (int i = threadIdx.x; i < 30; i += TILE * TPB) s_lf[i] = plom_lfact_tab[i];
    }
    if constexpr (sizeof(R) == 4 || V != 6 || MIF) __syncthreads();
    const int J = act ? f->J : 0, Jp = act ? f->Jp : 0;
    const int j = tile * TILE + tid;
    const bool valid = act && j < J;

    R X[M::S];
    int a = j;
    if (act) {
        if (n == 0 && !MIF) {
            const R *x0 = (const R *)f->x0;
#pragma unroll
            for (int k = 0; k < M::S; k++) X[k] = x0[k];
        } else {
            /* (with MIF, interval 0 starts from the particles' own initial states in X[0]) */
            if (n > 0) {
                if (valid && f->rflag[n - 1]) a = select_ancestor(f, j, n - 1);
                if (valid && f->ancestry) f->ancestry[(size_t)(n - 1) * J + j] = a;
            }
            const R *Xin = (const R *)f->X[n & 1];
            if (!valid) a = 0;
#pragma unroll
            for (int k = 0; k < M::S; k++) X[k] = Xin[(size_t)k * Jp + a];
        }
        if (n == 0 && f->allX && valid) {
            R *A = (R *)f->allX;
#pragma unroll
            for (int k = 0; k < M::S; k++) A[(size_t)k * Jp + j] = X[k];
        }
    }
    /* per-particle slots of pv (MIF only; the array is unused otherwise) */
    double lp[MIF ? M::NPP : 1];
    if constexpr (MIF) {
        if (valid) {
            if constexpr (sizeof(R) == 8) mif_particle<M::NPP>(f, n, j, a, s_bm, lp);
            else mif_particle<M::NPP>(f, n, j, a, plom_bm_tab, lp);
        }
    }

    const double t0 = act ? f->times[n] : 0.0, t1 = act ? f->times[n + 1] : 0.0;
    /* the time-only functions of the interval's substeps (the seasonal forcing ...): the same
     * for every particle of the tile, so each substep time is evaluated once, by one
     * thread, into the sub-block's LDS table; the generated step reads the table (identical
     * values: the same function of the same accumulated t), or evaluates them itself when
     * they do not fit */
    const R *tft = nullptr;
    if constexpr (M::NTF > 0 && V != 6) {
        __shared__ double s_tf_raw[PLOM_TF_CAP * TPB];
        R *s_tf = (R *)(s_tf_raw + q * PLOM_TF_CAP);
        const int K = act ? plom_substeps(t0, t1, f->dt) : 0;
        const bool fits = act && K * M::NTF <= PLOM_TF_CAP;
        if (fits) {
            for (int k = tid; k < K; k += TILE) {
                double t = t0;
                for (int r = 0; r < k; r++) t += f->dt;
                M::time_funcs((R)t, cov_steps(f, n) + (size_t)k * (size_t)f->cov_cs, s_tf + k * M::NTF);
            }
        }
        if constexpr (TPB == 1) {
            if (fits) __syncthreads();   /* uniform over the block */
        } else {
            __syncthreads();
        }
        if (fits) tft = s_tf;
    }
    if (valid) {
        if constexpr (V == 6) {
            /* the ode implementation: deterministic, no random numbers */
            const plom_spline_set *sp = f->spl_on ? &f->spl : nullptr;
            if constexpr (MIF) M::interval_ode_pp((double *)X, f->pv, lp, f->cov, sp, t0, t1, f->dt, f->eps_abs, f->eps_rel);
            else M::interval_ode((double *)X, f->pv, f->cov, sp, t0, t1, f->dt, f->eps_abs, f->eps_rel);
        } else if constexpr (sizeof(R) == 8) {
            plom_nstream64 ns;
            plom_nstream64_init(&ns, f->seed, f->stream, (uint32_t)n, (uint32_t)j, s_icdf);
            if constexpr (PSR) plom_psr_tables(&ns, s_bm, s_lf);
            if constexpr (MIF) M::template interval_pp<V>((double *)X, f->pv, lp, &ns, t0, t1, f->dt, f->sqdt, cov_steps(f, n), f->cov_cs, f->drift_on ? &f->dv : nullptr, (const double *)tft);
            else M::template interval<V>((double *)X, f->pv, &ns, t0, t1, f->dt, f->sqdt, cov_steps(f, n), f->cov_cs, f->drift_on ? &f->dv : nullptr, (const double *)tft);
        } else {
            plom_nstream32 ns;
            plom_nstream32_init(&ns, f->seed, f->stream, (uint32_t)n, (uint32_t)j, s_icdf32);
            if constexpr (MIF) {
                float lpf[M::NPP];
#pragma unroll
                for (int k = 0; k < M::NPP; k++) lpf[k] = (float)lp[k];
                M::template interval_pp<V>((float *)X, f->pvf, lpf, &ns, t0, t1, f->dt, (float)f->sqdt, cov_steps(f, n), f->cov_cs,
                                           f->drift_on ? &f->dv : nullptr, (const float *)tft);
            } else {
                M::template interval<V>((float *)X, f->pvf, &ns, t0, t1, f->dt, (float)f->sqdt, cov_steps(f, n), f->cov_cs,
                                        f->drift_on ? &f->dv : nullptr, (const float *)tft);
            }
        }
        R *Xout = (R *)f->X[(n + 1) & 1];
#pragma unroll
        for (int k = 0; k < M::S; k++) Xout[(size_t)k * Jp + j] = X[k];
        /* PLOM_FLAG_LINEAGE: after intervals K-1, 2K-1, ... (below N-1) a copy of the cloud,
         * where the segments of the lineage replay start */
        if (f->lin_snap && (n + 1) % f->lin_K == 0 && n + 1 < f->N) {
            R *A = (R *)f->lin_snap + (size_t)((n + 1) / f->lin_K - 1) * M::S * Jp;
#pragma unroll
            for (int k = 0; k < M::S; k++) A[(size_t)k * Jp + j] = X[k];
        }
        if (f->allX) {
            R *A = (R *)f->allX + (size_t)(n + 1) * M::S * Jp;
#pragma unroll
            for (int k = 0; k < M::S; k++) A[(size_t)k * Jp + j] = X[k];
        }
    }

    const bool wt = act && !(f->flags & PLOM_FLAG_NO_FILTER) && f->nonan[n] > 0;
    if constexpr (TPB == 1) {
        if (!wt) return;       /* uniform over the block */
    }
    double l = -PLOM_INF;
    if (valid && wt) {
        double obs[M::N_TS];
        int nf = 0;
        const double *cvd = cov_at(f, n + 1);
        if constexpr (MIF) {
            M::proj2obs_pp(X, f->pv, lp, cvd, obs);
            l = M::loglik_pp(obs, f->data + (size_t)n * f->n_ts, f->pv, lp, cvd, t1, f->like_min, &nf);
        } else {
            M::proj2obs(X, f->pv, cvd, obs);
            l = M::loglik(obs, f->data + (size_t)n * f->n_ts, f->pv, cvd, t1, f->like_min, &nf);
        }
        if (MIF && f->pm_prior) {
            /* mif --prior: the prior terms of the particle's theta for this interval (the
             * table holds what mif_particle wrote for it), then the failure test on the
             * patched weight */
            const double *th = f->th[(n + 1) & 1];
            double pt = 0.0;
            for (int o = 0; o < f->n_theta; o++)
                if (f->pm_prior[o] != PLOM_PRIOR_NONE && n < f->pm_puntil[o])
                    pt = pt + plom_mif_prior_term(f->pm_prior[o], f->pm_ppow[o], f->pm_tr[o], th[(size_t)o * Jp + j],
                                                  f->pm_min[o], f->pm_max[o], f->like_min);
            l = plom_mif_prior_patch(l, nf == M::N_TS, pt, f->all_fail_llike);
        } else if (nf == M::N_TS) {
            /* failure: every series observed and floored at LIKE_MIN (smc.c:44) */
            l = -PLOM_INF;
        }
    }
    if (wt) f->logw[j] = l;
    double *ql = lds + 4 * q;
    const double mt = max256(l, ql);
    double w = 0.0;
    if (mt != -PLOM_INF && l != -PLOM_INF) w = plom_exp(l - mt);
    const double s = tree256(w, ql);
    const double s2 = tree256(w * w, ql);
    if (wt && tid == 0) { f->tmax[tile] = mt; f->tsum[tile] = s; f->tsum2[tile] = s2; }
}

/* The lineage of an ode filter (one block of one wave; lane 0 works): the final particle by
 * inverse CDF on the last interval's cumulative weights (equal weights when that interval
 * had no data, was not filtered or every particle failed), its ancestors by the stored
 * ancestry, then its states regenerated from x0 with the same counter-based noise the
 * filter drew, interval by interval: s[N-1] = first i with C_i 2^33 >= (2x+1) Q;
 * s[n-1] = ancestry[n-1][s[n]]; state[n+1] = interval_n(state[n], noise(n, s[n])). */
template <class M, typename R, int V>
__global__ __launch_bounds__(64) void k_lineage(plom_dfilter *f)
{
    constexpr bool PSR = V == 4 || V == 5;
    __shared__ double s_icdf[PLOM_ICDF_TAB_SIZE];
    __shared__ float s_icdf32[PLOM_ICDF_TAB_SIZE];
    __shared__ double s_bm[PSR ? PLOM_BM_TAB_SIZE : 1];
    __shared__ double s_lf[PSR ? 30 : 1];
    if constexpr (PSR) {
        for (int i = threadIdx.x; i < PLOM_BM_TAB_SIZE; i += 64) s_bm[i] = plom_bm_tab[i];
        for (int i = threadIdx.x; i < 30; i += 64) s_lf[i] = plom_lfact_tab[i];
    }
    if constexpr (sizeof(R) == 8 && V != 6) {
        for (int i = threadIdx.x; i < PLOM_ICDF_TAB_SIZE; i += 64) s_icdf[i] = plom_icdf_tab[i];
        __syncthreads();
    }
    if constexpr (sizeof(R) == 4) {
        for (int i = threadIdx.x; i < PLOM_ICDF_TAB_SIZE; i += 64) s_icdf32[i] = plom_icdf_tab32[i];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const int N = f->N, J = f->J;
    if (N == 0 || !f->lin_slots) return;
    const uint64_t unum = plom_lineage_num(f->seed, f->stream);
    int j;
    if (!(f->flags & PLOM_FLAG_NO_FILTER) && f->nonan[N - 1] > 0 && f->rflag[N - 1]) {
        const uint64_t Q = f->tincl[f->T - 1];
        j = select_first(f, (plom_u128)unum * (plom_u128)Q, 1ull << 33);
    } else {
        j = (int)((unum * (uint64_t)J + (1ull << 33) - 1ull) >> 33) - 1;
    }
    j = j < 0 ? 0 : (j >= J ? J - 1 : j);
    f->lin_slots[N - 1] = j;
    for (int n = N - 1; n >= 1; n--) {
        j = f->ancestry[(size_t)(n - 1) * J + j];
        f->lin_slots[n - 1] = j;
    }
    R X[M::S];
    const R *x0 = (const R *)f->x0;
#pragma unroll
    for (int k = 0; k < M::S; k++) {
        X[k] = x0[k];
        f->lin_state[k] = (double)X[k];
    }
    for (int n = 0; n < N; n++) {
        const double t0 = f->times[n], t1 = f->times[n + 1];
        const uint32_t jn = (uint32_t)f->lin_slots[n];
        if constexpr (V == 6) {
            M::interval_ode((double *)X, f->pv, f->cov, f->spl_on ? &f->spl : nullptr, t0, t1, f->dt, f->eps_abs,
                            f->eps_rel);
        } else if constexpr (sizeof(R) == 8) {
            plom_nstream64 ns;
            plom_nstream64_init(&ns, f->seed, f->stream, (uint32_t)n, jn, s_icdf);
            if constexpr (PSR) plom_psr_tables(&ns, s_bm, s_lf);
            M::template interval<V>((double *)X, f->pv, &ns, t0, t1, f->dt, f->sqdt, cov_steps(f, n), f->cov_cs,
                                    f->drift_on ? &f->dv : nullptr, (const double *)nullptr);
        } else {
            plom_nstream32 ns;
            plom_nstream32_init(&ns, f->seed, f->stream, (uint32_t)n, jn, s_icdf32);
            M::template interval<V>((float *)X, f->pvf, &ns, t0, t1, f->dt, (float)f->sqdt, cov_steps(f, n), f->cov_cs,
                                    f->drift_on ? &f->dv : nullptr, (const float *)nullptr);
        }
#pragma unroll
        for (int k = 0; k < M::S; k++) f->lin_state[(size_t)(n + 1) * M::S + k] = (double)X[k];
    }
    /* observation mean and one drawn observation per row (print_p_X, print.c:312-341:
     * obs_mean and observation(), core_template.c:369-412: mean + sd N(0,1), clipped at 0) */
    for (int n = 0; n <= N; n++) {
        double xs[M::S], obs[M::N_TS], mu[M::N_TS], var[M::N_TS];
#pragma unroll
        for (int k = 0; k < M::S; k++) xs[k] = f->lin_state[(size_t)n * M::S + k];
        M::proj2obs(xs, f->pv, cov_at(f, n), obs);
        M::obs_moments(obs, f->pv, cov_at(f, n), f->times[n], mu, var);
        plom_lineage_obs(f->seed, f->stream, (uint32_t)n, M::N_TS, mu, var, plom_bm_tab, f->lin_obs + (size_t)n * M::N_TS * 2);
    }
}

/* The lineage of a filter of the sde and psr implementations, in three launches.
 *
 * k_lineage_slots (one lane): the final particle by inverse CDF on the last interval's
 * cumulative weights (equal weights when that interval had no data, was not filtered or
 * every particle failed), then its ancestors by the stored ancestry:
 * s[N-1] = first i with C_i 2^33 >= (2x+1) Q; s[n-1] = ancestry[n-1][s[n]].
 *
 * k_lineage_seg (one wave per segment, lane 0 works): the states regenerated with the same
 * counter-based noise the filter drew, interval by interval,
 * state[n+1] = interval_n(state[n], noise(n, s[n])).  Segment b replays intervals
 * [bK, min((b+1)K, N)); it starts from x0 (b = 0) or from the cloud the propagate kernel
 * copied after interval bK - 1, at particle s[bK - 1]: that copy is the state the
 * sequential replay reaches there (the same function of the same inputs), so the segments
 * give the sequential replay's states bit for bit.
 *
 * k_lineage_obs (one wave per row): the observation mean and one drawn observation of
 * every row. */
#define PLOM_LIN_SEG_MAX 16      /* segments of the lineage replay (grid of k_lineage_seg) */
#define PLOM_LIN_OBS_BLOCKS 1024 /* grid of k_lineage_obs: rows b, b + 1024, ... */
template <class M>
__global__ __launch_bounds__(64) void k_lineage_slots(plom_dfilter *f)
{
    if (threadIdx.x != 0) return;
    const int N = f->N, J = f->J;
    if (N == 0 || !f->lin_slots) return;
    const uint64_t unum = plom_lineage_num(f->seed, f->stream);
    int j;
    if (!(f->flags & PLOM_FLAG_NO_FILTER) && f->nonan[N - 1] > 0 && f->rflag[N - 1]) {
        const uint64_t Q = f->tincl[f->T - 1];
        j = select_first(f, (plom_u128)unum * (plom_u128)Q, 1ull << 33);
    } else {
        j = (int)((unum * (uint64_t)J + (1ull << 33) - 1ull) >> 33) - 1;
    }
    j = j < 0 ? 0 : (j >= J ? J - 1 : j);
    f->lin_slots[N - 1] = j;
    for (int n = N - 1; n >= 1; n--) {
        j = f->ancestry[(size_t)(n - 1) * J + j];
        f->lin_slots[n - 1] = j;
    }
}

template <class M, typename R, int V>
__global__ __launch_bounds__(64) void k_lineage_seg(plom_dfilter *f)
{
    constexpr bool PSR = V == 4 || V == 5;
    const int N = f->N;
    if (N == 0 || !f->lin_slots || (int)blockIdx.x >= f->lin_nseg) return;   /* the whole block */
    __shared__ double s_icdf[PLOM_ICDF_TAB_SIZE];
    __shared__ float s_icdf32[PLOM_ICDF_TAB_SIZE];
    __shared__ double s_bm[PSR ? PLOM_BM_TAB_SIZE : 1];
    __shared__ double s_lf[PSR ? 30 : 1];
    if constexpr (PSR) {
        for (int i = threadIdx.x; i < PLOM_BM_TAB_SIZE; i += 64) s_bm[i] = plom_bm_tab[i];
        for (int i = threadIdx.x; i < 30; i += 64) s_lf[i] = plom_lfact_tab[i];
    }
    if constexpr (sizeof(R) == 8) {
        for (int i = threadIdx.x; i < PLOM_ICDF_TAB_SIZE; i += 64) s_icdf[i] = plom_icdf_tab[i];
    }
    if constexpr (sizeof(R) == 4) {
        for (int i = threadIdx.x; i < PLOM_ICDF_TAB_SIZE; i += 64) s_icdf32[i] = plom_icdf_tab32[i];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int K = f->lin_K, b = (int)blockIdx.x;
    const int n0 = b * K, n1 = (b + 1) * K < N ? (b + 1) * K : N;
    R X[M::S];
    if (b == 0) {
        const R *x0 = (const R *)f->x0;
#pragma unroll
        for (int k = 0; k < M::S; k++) {
            X[k] = x0[k];
            f->lin_state[k] = (double)X[k];
        }
    } else {
        const R *A = (const R *)f->lin_snap + (size_t)(b - 1) * M::S * f->Jp;
        const size_t a = (size_t)f->lin_slots[n0 - 1];
#pragma unroll
        for (int k = 0; k < M::S; k++) X[k] = A[(size_t)k * f->Jp + a];
    }
    for (int n = n0; n < n1; n++) {
        const double t0 = f->times[n], t1 = f->times[n + 1];
        const uint32_t jn = (uint32_t)f->lin_slots[n];
        if constexpr (sizeof(R) == 8) {
            plom_nstream64 ns;
            plom_nstream64_init(&ns, f->seed, f->stream, (uint32_t)n, jn, s_icdf);
            if constexpr (PSR) plom_psr_tables(&ns, s_bm, s_lf);
            M::template interval<V>((double *)X, f->pv, &ns, t0, t1, f->dt, f->sqdt, cov_steps(f, n), f->cov_cs,
                                    f->drift_on ? &f->dv : nullptr, (const double *)nullptr);
        } else {
            plom_nstream32 ns;
            plom_nstream32_init(&ns, f->seed, f->stream, (uint32_t)n, jn, s_icdf32);
            M::template interval<V>((float *)X, f->pvf, &ns, t0, t1, f->dt, (float)f->sqdt, cov_steps(f, n), f->cov_cs,
                                    f->drift_on ? &f->dv : nullptr, (const float *)nullptr);
        }
#pragma unroll
        for (int k = 0; k < M::S; k++) f->lin_state[(size_t)(n + 1) * M::S + k] = (double)X[k];
    }
}

/* observation mean and one drawn observation per row (print_p_X, print.c:312-341:
 * obs_mean and observation(), core_template.c:369-412: mean + sd N(0,1), clipped at 0) */
template <class M>
__global__ __launch_bounds__(64) void k_lineage_obs(plom_dfilter *f)
{
    if (threadIdx.x != 0 || !f->lin_slots) return;
    const int N = f->N;
    for (int n = (int)blockIdx.x; n <= N; n += (int)gridDim.x) {
        double xs[M::S], obs[M::N_TS], mu[M::N_TS], var[M::N_TS];
#pragma unroll
        for (int k = 0; k < M::S; k++) xs[k] = f->lin_state[(size_t)n * M::S + k];
        M::proj2obs(xs, f->pv, cov_at(f, n), obs);
        M::obs_moments(obs, f->pv, cov_at(f, n), f->times[n], mu, var);
        plom_lineage_obs(f->seed, f->stream, (uint32_t)n, M::N_TS, mu, var, plom_bm_tab, f->lin_obs + (size_t)n * M::N_TS * 2);
    }
}

/* ------------------------------------------------------------------ hat and residuals */
/* The summaries of interval n (launched after its weights and scan, before the next
 * propagate reads the cloud).  Weights are the fixed-point integers of the resampling,
 * q_j = floor(exp(l_j - m)/S 2^62) (0 for failed particles), or 1 for every particle when
 * the interval was not weighted; Q is their total.
 *   mean     tree sum over tiles of the tile trees of value x q, divided by Q;
 *   low95    the smallest value v with 40 W(<= v) >= Q, W(<= v) the total weight of the
 *            particles whose value is at most v;
 *   high95   the smallest v with 40 W(<= v) >= 39 Q
 * (hat.c:29-55: the first sorted value at which the cumulated weight reaches 2.5 % /
 * 97.5 %; exact in integers here, and independent of the order of equal values).
 * Each tile sorts its values (bitonic, in LDS) with their weights and keeps the inclusive
 * weight prefix; one block per quantity and bound then bisects the 64-bit key space,
 * summing the tiles' prefixes exactly. */
__device__ inline uint64_t hat_key(double x)
{
    const uint64_t u = plom_d2u(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ inline double hat_unkey(uint64_t k)
{
    return plom_u2d((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k);
}

__device__ inline bool hat_weighted(const plom_dfilter *f, int n)
{
    return !(f->flags & PLOM_FLAG_NO_FILTER) && f->nonan[n] > 0 && f->rflag[n];
}

/* number of systematic points k in [0, J) with (k 2^33 + u) Q <= C J 2^33, i.e. of
 * particles whose ancestor index is at most the particle with inclusive prefix C */
__device__ inline uint64_t hat_count_le(uint64_t C, uint64_t J, uint64_t unum, uint64_t Q)
{
    const plom_u128 X = (plom_u128)C * J << 33;
    const double est = (double)C * (double)J / (double)Q - (double)unum * 0x1p-33;
    int64_t k = est < 0.0 ? 0 : (int64_t)est + 1;
    if (k > (int64_t)J) k = (int64_t)J;
    while (k > 0 && (plom_u128)(((uint64_t)(k - 1) << 33) + unum) * Q > X) k--;
    while (k < (int64_t)J && (plom_u128)(((uint64_t)k << 33) + unum) * Q <= X) k++;
    return (uint64_t)k;
}

/* the copies o_j of particle j in the resampled cloud of interval n (1 each when the
 * interval did not resample): the systematic points between its exclusive and inclusive
 * prefix */
__device__ inline uint64_t hat_offspring(const plom_dfilter *f, int n, int j, uint64_t unum, uint64_t Q)
{
    if (!f->rflag[n]) return 1;
    const int tile = j / TILE;
    const uint64_t base = tile ? f->tincl[tile - 1] : 0ull;
    const uint64_t C = base + f->cum[j];
    const uint64_t Cm = (j % TILE) ? base + f->cum[j - 1] : base;
    return hat_count_le(C, (uint64_t)f->J, unum, Q) - hat_count_le(Cm, (uint64_t)f->J, unum, Q);
}

/* pairwise tree (contiguous, zero padded to a power of two) over the T tile values of a
 * filter: per chunk of 256 tiles tree256, then over the chunks (at most 32) by thread 0;
 * every thread gets the total */
__device__ inline double tiles_tree(const double *v, int T, double *lds, double *chunks)
{
    const int nch = (T + TILE - 1) / TILE;
    for (int c = 0; c < nch; c++) {
        const int t = c * TILE + threadIdx.x;
        const double s = tree256(t < T ? v[t] : 0.0, lds);
        if (threadIdx.x == 0) chunks[c] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int P = 1;
        while (P < nch) P <<= 1;
        for (int c = nch; c < P; c++) chunks[c] = 0.0;
        for (int h = 1; h < P; h <<= 1)
            for (int i = 0; i < P; i += 2 * h) chunks[i] = chunks[i] + chunks[i + h];
        lds[4] = chunks[0];
    }
    __syncthreads();
    const double out = lds[4];
    __syncthreads();
    return out;
}

/* the (count, mean, M2) of the T tiles merged along the fixed pairwise tree (within a chunk
 * of 256 tiles, then over the chunks padded with empty sets to a power of two): the M2 of
 * the union; every thread gets it */
__device__ inline double tiles_chan(const double *w, const double *m, const double *q, int T, double *sc, double *sm,
                                    double *sq, double *cc, double *cm, double *cq)
{
    const int nch = (T + TILE - 1) / TILE;
    for (int c = 0; c < nch; c++) {
        const int t = c * TILE + threadIdx.x;
        sc[threadIdx.x] = t < T ? w[t] : 0.0;
        sm[threadIdx.x] = t < T ? m[t] : 0.0;
        sq[threadIdx.x] = t < T ? q[t] : 0.0;
        __syncthreads();
        for (int h = 1; h < TILE; h <<= 1) {
            const int i = threadIdx.x * 2 * h;
            double c2 = 0.0, m2 = 0.0, q2 = 0.0;
            if (i < TILE) plom_mif_combine(sc[i], sm[i], sq[i], sc[i + h], sm[i + h], sq[i + h], &c2, &m2, &q2);
            __syncthreads();
            if (i < TILE) { sc[i] = c2; sm[i] = m2; sq[i] = q2; }
            __syncthreads();
        }
        if (threadIdx.x == 0) { cc[c] = sc[0]; cm[c] = sm[0]; cq[c] = sq[0]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int P = 1;
        while (P < nch) P <<= 1;
        for (int c = nch; c < P; c++) { cc[c] = 0.0; cm[c] = 0.0; cq[c] = 0.0; }
        for (int h = 1; h < P; h <<= 1)
            for (int i = 0; i < P; i += 2 * h)
                plom_mif_combine(cc[i], cm[i], cq[i], cc[i + h], cm[i + h], cq[i + h], &cc[i], &cm[i], &cq[i]);
        sq[0] = cq[0];
    }
    __syncthreads();
    const double out = sq[0];
    __syncthreads();
    return out;
}

/* Block (tile, y), y < PLOM_HAT_GROUPS, sorts the quantities [y QB, (y + 1) QB) of its
 * tile, QB = the quantities over PLOM_HAT_GROUPS blocks; block y = PLOM_HAT_GROUPS sums the
 * tile's parts of the residuals. */
#define PLOM_HAT_GROUPS 8
template <class M, typename R>
__global__ __launch_bounds__(TILE) void k_hat_tiles(plom_dfilter *f)
{
    __shared__ double lds[8];
    __shared__ uint64_t lds64[8];
    __shared__ uint64_t s_key[TILE];
    __shared__ uint64_t s_w[TILE];
    const int n = f->step - 1;
    if (n < 0 || n >= f->N || !f->hkey) return;
    const int J = f->J, Jp = f->Jp, T = f->T, tile = blockIdx.x;
    const int j = tile * TILE + threadIdx.x;
    const bool valid = j < J;
    double v[M::N_HAT];
    double xs[M::S];
    uint64_t q = 0;
    if (valid) {
        const R *Xs = (const R *)f->X[(n + 1) & 1];
#pragma unroll
        for (int k = 0; k < M::S; k++) xs[k] = (double)Xs[(size_t)k * Jp + j];
        M::hat_values(xs, f->pv, cov_at(f, n + 1), f->drift_on ? &f->dv : nullptr, f->times[n + 1], v);
        if (hat_weighted(f, n)) {
            const double l = f->logw[j];
            if (l != -PLOM_INF) q = (uint64_t)((plom_exp(l - f->scal[0]) / f->scal[1]) * 0x1p62);
        } else {
            q = 1;
        }
    }
    constexpr int QB = (M::N_HAT + PLOM_HAT_GROUPS - 1) / PLOM_HAT_GROUPS;
    const int k0 = (int)blockIdx.y * QB, k1 = k0 + QB < M::N_HAT ? k0 + QB : M::N_HAT;
    for (int k = k0; k < k1; k++) {
        const double part = tree256(valid ? v[k] * (double)q : 0.0, lds);
        if (threadIdx.x == 0) f->hpart[(size_t)k * T + tile] = part;
        s_key[threadIdx.x] = valid ? hat_key(v[k]) : ~0ull;
        s_w[threadIdx.x] = q;
        __syncthreads();
        for (int size = 2; size <= TILE; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                const int t = threadIdx.x, p = t ^ stride;
                if (p > t) {
                    const bool up = (t & size) == 0;
                    const uint64_t a = s_key[t], b = s_key[p];
                    if ((a > b) == up) {
                        s_key[t] = b; s_key[p] = a;
                        const uint64_t wa = s_w[t];
                        s_w[t] = s_w[p]; s_w[p] = wa;
                    }
                }
                __syncthreads();
            }
        }
        const uint64_t key = s_key[threadIdx.x];
        const uint64_t pref = scan256(s_w[threadIdx.x], lds64);
        f->hkey[(size_t)k * Jp + j] = key;
        f->hpref[(size_t)k * Jp + j] = pref;
        __syncthreads();
    }
    /* the tile's part of the prediction residuals (print.c:632-701) over the resampled
     * cloud: with o_j the copies of particle j, W = sum o, and per series the sums
     * A = sum o mu and B = sum o var (tree256), the tile mean A / W and
     * Q = sum o (mu - A / W)^2; k_pred_res merges the tiles */
    if (blockIdx.y != PLOM_HAT_GROUPS || !f->prw || (f->flags & PLOM_FLAG_NO_FILTER) || f->nonan[n] == 0) return;   /* uniform */
    const bool res_on = f->rflag[n] != 0;
    const uint64_t Q = res_on ? f->tincl[T - 1] : 0;
    const uint64_t unum = plom_resample_num(f->seed, f->stream, (uint32_t)n);
    uint64_t o = 0;
    double mu[M::N_TS], var[M::N_TS];
#pragma unroll
    for (int ts = 0; ts < M::N_TS; ts++) { mu[ts] = 0.0; var[ts] = 0.0; }
    if (valid) {
        o = hat_offspring(f, n, j, unum, Q);
        if (o) {
            double obs[M::N_TS];
            M::proj2obs(xs, f->pv, cov_at(f, n + 1), obs);
            M::obs_moments(obs, f->pv, cov_at(f, n + 1), f->times[n + 1], mu, var);
        }
    }
    const double od = (double)o;
    const double W = tree256(od, lds);
    if (threadIdx.x == 0) f->prw[tile] = W;
    double *pp = f->prp;
#pragma unroll
    for (int ts = 0; ts < M::N_TS; ts++) {
        const double A = tree256(o ? od * mu[ts] : 0.0, lds);
        const double B = tree256(o ? od * var[ts] : 0.0, lds);
        const double mt = W > 0.0 ? A / W : 0.0;
        const double d = mu[ts] - mt;
        const double Qs = tree256(o ? od * (d * d) : 0.0, lds);
        if (threadIdx.x == 0) {
            pp[(size_t)(0 * M::N_TS + ts) * T + tile] = A;
            pp[(size_t)(1 * M::N_TS + ts) * T + tile] = B;
            pp[(size_t)(2 * M::N_TS + ts) * T + tile] = mt;
            pp[(size_t)(3 * M::N_TS + ts) * T + tile] = Qs;
        }
    }
}


/* prediction residuals of interval n over the resampled cloud (print.c:632-701), one block
 * per series from the tiles' parts (k_hat_tiles): pred = sum_t A_t / J, var_obs =
 * sum_t B_t / J (pairwise trees over the tiles), var_state = M2 / (J - 1) with M2 the
 * (count, mean, M2) merge of the tiles' (W_t, A_t / W_t, Q_t) along the fixed tree, that is
 * sum o_j (mu_j - pred)^2; res = (y - pred) / sqrt(var_state + var_obs) */
template <class M, typename R>
__global__ __launch_bounds__(TILE) void k_pred_res(plom_dfilter *f)
{
    __shared__ double lds[8];
    __shared__ double chunks[MAX_TILES / TILE + 1];
    __shared__ double sc[TILE], sm[TILE], sq[TILE];
    __shared__ double cc[MAX_TILES / TILE + 1], cm[MAX_TILES / TILE + 1], cq[MAX_TILES / TILE + 1];
    const int n = f->step - 1;
    const int ts = blockIdx.x;
    if (n < 0 || n >= f->N || !f->pres || ts >= M::N_TS) return;
    double *out = f->pres + (size_t)n * M::N_TS * 2;
    if ((f->flags & PLOM_FLAG_NO_FILTER) || f->nonan[n] == 0) {
        if (threadIdx.x < 2) out[2 * ts + threadIdx.x] = PLOM_NAN;
        return;
    }
    const int J = f->J, T = f->T;
    const double *pp = f->prp;
    const double pred = tiles_tree(pp + (size_t)(0 * M::N_TS + ts) * T, T, lds, chunks) / (double)J;
    const double vobs = tiles_tree(pp + (size_t)(1 * M::N_TS + ts) * T, T, lds, chunks) / (double)J;
    const double m2 = tiles_chan(f->prw, pp + (size_t)(2 * M::N_TS + ts) * T, pp + (size_t)(3 * M::N_TS + ts) * T, T,
                                 sc, sm, sq, cc, cm, cq);
    if (threadIdx.x == 0) {
        const double vstate = m2 / (double)(J - 1);
        const double y = f->data[(size_t)n * M::N_TS + ts];
        out[2 * ts] = pred;
        out[2 * ts + 1] = (y - pred) / plom_sqrt(vstate + vobs);
    }
}

/* row 0 of the hat: every particle starts at x0 */
template <class M, typename R>
__global__ void k_hat0(plom_dfilter *f)
{
    if (threadIdx.x != 0 || !f->hat) return;
    const R *x0 = (const R *)f->x0;
    double xs[M::S], v[M::N_HAT];
#pragma unroll
    for (int k = 0; k < M::S; k++) xs[k] = (double)x0[k];
    M::hat_values(xs, f->pv, cov_at(f, 0), f->drift_on ? &f->dv : nullptr, f->times[0], v);
    for (int k = 0; k < M::N_HAT; k++) f->hat[3 * k] = f->hat[3 * k + 1] = f->hat[3 * k + 2] = v[k];
}

/* ------------------------------------------------------------------ model table */
/* What a model's translation unit exports: its dimensions and the host launchers of its
 * kernels.  abi and dfilter_bytes let the library refuse a model library compiled against
 * another version of this header (plom_model_load). */
#define PLOM_MODEL_ABI 6
struct model_entry {
    int32_t abi;            /* PLOM_MODEL_ABI */
    int32_t dfilter_bytes;  /* sizeof(plom_dfilter) */
    const char *name;
    plom_model_info info;
    /* one propagate launch over ntiles tiles, tpb (1, 2 or 4) tiles per block */
    void (*launch_prop)(int prec, int variant, int mif, int ntiles, int tpb, hipStream_t s, plom_dfilter *df,
                        const int2 *tiles);
    void (*launch_lineage)(int prec, int variant, hipStream_t s, plom_dfilter *df);
    void (*launch_hat)(int prec, hipStream_t s, plom_dfilter *df, int ntiles, int n_hat);
    void (*launch_hat0)(int prec, hipStream_t s, plom_dfilter *df);
    /* host builds of the generated observation functions (stored clouds, read on the host) */
    void (*proj2obs_h)(const double *X, const double *pvd, const double *cvd, double *obs);
    void (*moments_h)(const double *obs, const double *pvd, const double *cvd, double t, double *mean, double *var);
};


/* one propagate launch over ntiles tiles (tile-map entries, or tiles of one filter without
 * a map), tpb tiles per block */
template <class M, typename R, int V, bool MIF, int TPB>
static void launch_prop_k(int ntiles, hipStream_t s, plom_dfilter *df, const int2 *tiles)
{
    hipLaunchKernelGGL((k_propagate<M, R, V, MIF, TPB>), dim3((ntiles + TPB - 1) / TPB), dim3(TILE * TPB), 0, s, df,
                       tiles, ntiles);
}

/* tiles per block: 1, 2 or 4 for the sde and psr variants without per-particle parameters,
 * 1 otherwise (other values fall back to 1) */
template <class M, typename R, int V, bool MIF>
static void launch_prop_v(int ntiles, int tpb, hipStream_t s, plom_dfilter *df, const int2 *tiles)
{
    if constexpr (!MIF && V != 6) {
        if (tpb == 2) { launch_prop_k<M, R, V, MIF, 2>(ntiles, s, df, tiles); return; }
        if (tpb == 4) { launch_prop_k<M, R, V, MIF, 4>(ntiles, s, df, tiles); return; }
    }
    launch_prop_k<M, R, V, MIF, 1>(ntiles, s, df, tiles);
}

/* variants 0-3: SDE noise switches; 4, 5: psr with / without white noise; 6: ode (FP64 only) */
template <class M, typename R, bool MIF>
static void launch_prop_r(int variant, int ntiles, int tpb, hipStream_t s, plom_dfilter *df, const int2 *tiles)
{
    if constexpr (sizeof(R) == 8) {
        if (variant == 4) { launch_prop_v<M, R, 4, MIF>(ntiles, tpb, s, df, tiles); return; }
        if (variant == 5) { launch_prop_v<M, R, 5, MIF>(ntiles, tpb, s, df, tiles); return; }
        if (variant == 6) { launch_prop_v<M, R, 6, MIF>(ntiles, tpb, s, df, tiles); return; }
    }
    switch (variant) {
    case 0: launch_prop_v<M, R, 0, MIF>(ntiles, tpb, s, df, tiles); break;
    case 1: launch_prop_v<M, R, 1, MIF>(ntiles, tpb, s, df, tiles); break;
    case 2: launch_prop_v<M, R, 2, MIF>(ntiles, tpb, s, df, tiles); break;
    default: launch_prop_v<M, R, 3, MIF>(ntiles, tpb, s, df, tiles); break;
    }
}

template <class M>
static void launch_prop(int prec, int variant, int mif, int ntiles, int tpb, hipStream_t s, plom_dfilter *df,
                        const int2 *tiles)
{
    if (prec == PLOM_FP32) {
        if (mif) launch_prop_r<M, float, true>(variant, ntiles, tpb, s, df, tiles);
        else launch_prop_r<M, float, false>(variant, ntiles, tpb, s, df, tiles);
    } else {
        if (mif) launch_prop_r<M, double, true>(variant, ntiles, tpb, s, df, tiles);
        else launch_prop_r<M, double, false>(variant, ntiles, tpb, s, df, tiles);
    }
}

template <class M, typename R, int V>
static void launch_lineage_v(hipStream_t s, plom_dfilter *df)
{
    if constexpr (V == 6) {
        hipLaunchKernelGGL((k_lineage<M, R, V>), dim3(1), dim3(64), 0, s, df);
    } else {
        hipLaunchKernelGGL((k_lineage_slots<M>), dim3(1), dim3(64), 0, s, df);
        hipLaunchKernelGGL((k_lineage_seg<M, R, V>), dim3(PLOM_LIN_SEG_MAX), dim3(64), 0, s, df);
        hipLaunchKernelGGL((k_lineage_obs<M>), dim3(PLOM_LIN_OBS_BLOCKS), dim3(64), 0, s, df);
    }
}

template <class M, typename R>
static void launch_lineage_r(int variant, hipStream_t s, plom_dfilter *df)
{
    if constexpr (sizeof(R) == 8) {
        if (variant == 4) { launch_lineage_v<M, R, 4>(s, df); return; }
        if (variant == 5) { launch_lineage_v<M, R, 5>(s, df); return; }
        if (variant == 6) { launch_lineage_v<M, R, 6>(s, df); return; }
    }
    switch (variant) {
    case 0: launch_lineage_v<M, R, 0>(s, df); break;
    case 1: launch_lineage_v<M, R, 1>(s, df); break;
    case 2: launch_lineage_v<M, R, 2>(s, df); break;
    default: launch_lineage_v<M, R, 3>(s, df); break;
    }
}

template <class M>
static void launch_lineage(int prec, int variant, hipStream_t s, plom_dfilter *df)
{
    if (prec == PLOM_FP32) launch_lineage_r<M, float>(variant, s, df);
    else launch_lineage_r<M, double>(variant, s, df);
}

/* the per-tile hat sort and the residuals (the library then launches k_hat_mean and
 * k_hat_bounds) */
template <class M>
static void launch_hat(int prec, hipStream_t s, plom_dfilter *df, int ntiles, int n_hat)
{
    if (prec == PLOM_FP32) {
        hipLaunchKernelGGL((k_hat_tiles<M, float>), dim3(ntiles, PLOM_HAT_GROUPS + 1), dim3(TILE), 0, s, df);
        hipLaunchKernelGGL((k_pred_res<M, float>), dim3(M::N_TS), dim3(TILE), 0, s, df);
    } else {
        hipLaunchKernelGGL((k_hat_tiles<M, double>), dim3(ntiles, PLOM_HAT_GROUPS + 1), dim3(TILE), 0, s, df);
        hipLaunchKernelGGL((k_pred_res<M, double>), dim3(M::N_TS), dim3(TILE), 0, s, df);
    }
    (void)ntiles; (void)n_hat;
}

template <class M>
static void launch_hat0(int prec, hipStream_t s, plom_dfilter *df)
{
    if (prec == PLOM_FP32) hipLaunchKernelGGL((k_hat0<M, float>), dim3(1), dim3(64), 0, s, df);
    else hipLaunchKernelGGL((k_hat0<M, double>), dim3(1), dim3(64), 0, s, df);
}

#define PLOM_CAT_(a, b) a##b
#define PLOM_CAT(a, b) PLOM_CAT_(a, b)
#define PLOM_STR_(a) #a
#define PLOM_STR(a) PLOM_STR_(a)

/* the exported entry of model `lname` (generated constants UNAME_*, generated struct
 * plom_model_<lname>, host functions plom_<lname>_*) */
#define PLOM_DEFINE_MODEL(lname, UNAME)                                                              \
    extern "C" const model_entry *PLOM_CAT(plom_model_entry_, lname)(void)                           \
    {                                                                                                \
        static const model_entry e = {                                                               \
            PLOM_MODEL_ABI, (int32_t)sizeof(plom_dfilter), PLOM_STR(lname),                          \
            { PLOM_CAT(UNAME, _SIZE_PROJ), PLOM_CAT(UNAME, _N_TS), PLOM_CAT(UNAME, _N_CAC),            \
              PLOM_CAT(UNAME, _N_PAR_SV), PLOM_CAT(UNAME, _INC_OFFSET), PLOM_CAT(UNAME, _N_TS_INC_UNIQUE), \
              PLOM_CAT(UNAME, _PV_SIZE), PLOM_CAT(UNAME, _N_DEM), PLOM_CAT(UNAME, _N_ENV),             \
              PLOM_CAT(UNAME, _N_HAT), PLOM_CAT(UNAME, _N_PP), PLOM_CAT(UNAME, _FIX_OFF),              \
              PLOM_CAT(UNAME, _N_FIX), PLOM_CAT(UNAME, _N_DRIFT), PLOM_CAT(UNAME, _DRIFT_OFFSET) },     \
            launch_prop<PLOM_CAT(plom_model_, lname)>, launch_lineage<PLOM_CAT(plom_model_, lname)>,     \
            launch_hat<PLOM_CAT(plom_model_, lname)>, launch_hat0<PLOM_CAT(plom_model_, lname)>,         \
            PLOM_CAT(PLOM_CAT(plom_, lname), _proj2obs_f64), PLOM_CAT(PLOM_CAT(plom_, lname), _obs_moments) }; \
        return &e;                                                                                   \
    }

#endif /* PLOM_DEVICE_H */
