/*
 * plom_smc.hip -- the PLOM SMC particle filter on MI355X (gfx950): the model-independent
 * kernels and the C ABI (filters, batches, the server).
 *
 * One observation interval n of a filter (or of every filter of a batch, over a tile map
 * (filter, tile)) is four stream-ordered launches over tiles of 256 particles:
 *
 *   k_propagate<Model, Real, Variant, MIF>   (plom_device.h, one translation unit per model)
 *       select the ancestor of particle j from the previous interval's fixed-point
 *       cumulative weights (systematic resampling, smc.c:90-112), gather its state from the
 *       SoA cloud [S][Jp] (resample_X, smc.c:142-161), run the interval's Euler-Maruyama
 *       substeps in registers (generated step, core_template.c:215-339, driven as in
 *       prediction_util.c:275-287), the log-weight l_j (likelihood.c:56-92), and per tile
 *       the max m_t and the tree sums of exp(l - m_t) and of its square;
 *   k_weights   one block per filter: the tiles reduced to m, S, S2; loglik_n = m +
 *               log(S/J), ESS = S^2/S2, the all-fail branch (smc.c:27-83);
 *   k_scan      q_j = trunc(exp(l_j - m)/S 2^62), inclusive integer scan within the tile;
 *   k_advance   one block per filter: inclusive prefix of the tile totals, interval + 1.
 *
 * Canonical arithmetic shared with the oracle (oracle/smc_oracle.c): the sums over 256
 * particles and over tiles are the recursive contiguous pairwise tree (zero padded to a
 * power of two); the resampling compares (k 2^33 + 2u + 1) Q with C_i J 2^33 exactly in
 * 128-bit integers, C the integer prefix of q, Q its total, u the interval's 32-bit
 * uniform; select[k] is the first i with C_i J 2^33 >= (k 2^33 + 2u + 1) Q.
 *
 * Every filter keeps its interval counter in device memory, so a launch sequence can be
 * captured in a hipGraph and replayed.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>
#include <vector>
#include <algorithm>
#include <new>

#include <dlfcn.h>
#include <atomic>
#include <mutex>

#include "plom_device.h"
#include "plom_model_list.h"

/* ------------------------------------------------------------------ errors */
static thread_local char g_err[512];

static int set_err(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return set_err(e_ == hipErrorOutOfMemory ? PLOM_ERR_NOMEM : PLOM_ERR_HIP,  \
                           "%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
    } while (0)

extern "C" const char *plom_last_error(void) { return g_err; }

/* global reduction of the tile statistics by the last block of interval n */
__device__ void finalize_weights(plom_dfilter *f, int n, double *lds)
{
    __shared__ double s_chunk[MAX_TILES / TILE + 1][2];
    const int T = f->T;
    double m = -PLOM_INF;
    for (int t = threadIdx.x; t < T; t += TILE) m = fmax(m, f->tmax[t]);
    m = max256(m, lds);
    double llike_n, ess;
    int rflag;
    if (m == -PLOM_INF) {
        /* every particle failed: equal weights, identity ancestry (smc.c:55-70) */
        llike_n = f->all_fail_llike;
        ess = 0.0;
        rflag = 0;
    } else {
        const int nch = (T + TILE - 1) / TILE;
        for (int c = 0; c < nch; c++) {
            const int t = c * TILE + threadIdx.x;
            double v = 0.0, v2 = 0.0;
            if (t < T) {
                const double mt = f->tmax[t];
                if (mt != -PLOM_INF) {
                    const double e = plom_exp(mt - m);
                    v = f->tsum[t] * e;
                    v2 = (f->tsum2[t] * e) * e;
                }
            }
            v = tree256(v, lds);
            v2 = tree256(v2, lds);
            if (threadIdx.x == 0) { s_chunk[c][0] = v; s_chunk[c][1] = v2; }
        }
        __syncthreads();
        double S = 0.0, S2 = 0.0;
        if (threadIdx.x == 0) {
            int P = 1;
            while (P < nch) P <<= 1;
            for (int c = nch; c < P; c++) { s_chunk[c][0] = 0.0; s_chunk[c][1] = 0.0; }
            for (int h = 1; h < P; h <<= 1)
                for (int i = 0; i < P; i += 2 * h) {
                    s_chunk[i][0] = s_chunk[i][0] + s_chunk[i + h][0];
                    s_chunk[i][1] = s_chunk[i][1] + s_chunk[i + h][1];
                }
            S = s_chunk[0][0];
            S2 = s_chunk[0][1];
        }
        llike_n = m + plom_log(S / (double)f->J);
        ess = (S * S) / S2;
        rflag = 1;
        if (threadIdx.x == 0) { f->scal[0] = m; f->scal[1] = S; f->scal[2] = S2; }
    }
    if (threadIdx.x == 0) {
        if (!rflag) f->n_all_fail[0] += 1;
        f->llike_n[n] = llike_n;
        f->ess_n[n] = ess;
        f->rflag[n] = rflag;
        f->scal[3] = f->scal[3] + llike_n;
    }
}


/* one block per filter: the tiles' statistics -> loglik_n, ESS, resampling flag */
__global__ __launch_bounds__(TILE) void k_weights(plom_dfilter *filters, const int *fidx)
{
    __shared__ double lds[8];
    const int fi = fidx ? fidx[blockIdx.x] : (int)blockIdx.x;
    plom_dfilter *f = &filters[fi];
    const int n = f->step;
    if (n >= f->N) return;
    if ((f->flags & PLOM_FLAG_NO_FILTER) || f->nonan[n] == 0) return;
    finalize_weights(f, n, lds);
}

/* PLOM_FLAG_MIF, interval n with data, after k_scan: per tile, the sum of the weights
 * w_j = exp(l_j - m) (1 each when every particle failed, smc.c:55-70) and, per component,
 * the sum of w_j theta_j, the tile mean and the sum of squared deviations from it.  Every
 * sum is the tree of tree256 (a butterfly within each wave, then (w0 + w1) + (w2 + w3)).
 * Block (tile, y) takes components [y MIF_COMPS, (y + 1) MIF_COMPS): the grid spreads the
 * components over many blocks (the sums are chains of dependent shuffles, so the blocks'
 * number hides their latency).  Block y = 0 sums and writes the tile's weights.  The
 * interval n is an argument and the weights come from k_scan's copy (mw), not from the
 * log-weights and the maximum: so the kernel can run beside the next interval's
 * propagate, which overwrites those. */
#define MIF_COMPS 1
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) v = v + __shfl_xor(v, h, 64);
    return v;
}

__global__ __launch_bounds__(TILE) void k_mif_tiles(plom_dfilter *filters, const int2 *tiles, int n)
{
    __shared__ double lds[8];
    __shared__ double red[MIF_COMPS][4][2];
    __shared__ double smean[MIF_COMPS];
    int fi = 0, tile = blockIdx.x;
    if (tiles) { const int2 tt = tiles[blockIdx.x]; fi = tt.x; tile = tt.y; }
    plom_dfilter *f = &filters[fi];
    if (n >= f->N || !(f->flags & PLOM_FLAG_MIF) || f->nonan[n] == 0) return;
    const int J = f->J, Jp = f->Jp, T = f->T, nth = f->n_theta;
    const int o0 = (int)blockIdx.y * MIF_COMPS;
    if (o0 >= nth) return;                 /* uniform over the block */
    const int nc = nth - o0 < MIF_COMPS ? nth - o0 : MIF_COMPS;
    const int j = tile * TILE + threadIdx.x;
    const bool valid = j < J;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double w = 0.0;
    if (valid) w = f->rflag[n] ? f->mw[j] : 1.0;
    const double sw = blockIdx.y == 0 ? tree256(w, lds) : 0.0;   /* uniform over the block */
    const int rem = J - tile * TILE;
    const double ct = (double)(rem < TILE ? rem : TILE);
    const double *th = f->th[(n + 1) & 1] + (size_t)o0 * Jp;
    double v[MIF_COMPS];
#pragma unroll
    for (int c = 0; c < MIF_COMPS; c++) v[c] = (valid && c < nc) ? th[(size_t)c * Jp + j] : 0.0;
#pragma unroll
    for (int c = 0; c < MIF_COMPS; c++) {
        if (c < nc) {
            const double a = wave_sum(valid ? w * v[c] : 0.0);
            const double b = wave_sum(v[c]);
            if (lane == 0) { red[c][wv][0] = a; red[c][wv][1] = b; }
        }
    }
    __syncthreads();
    if (threadIdx.x < nc) {
        const int c = threadIdx.x;
        const double s1 = (red[c][0][0] + red[c][1][0]) + (red[c][2][0] + red[c][3][0]);
        const double mt = ((red[c][0][1] + red[c][1][1]) + (red[c][2][1] + red[c][3][1])) / ct;
        smean[c] = mt;
        f->mps[(size_t)(o0 + c) * T + tile] = s1;
        f->mpm[(size_t)(o0 + c) * T + tile] = mt;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MIF_COMPS; c++) {
        if (c < nc) {
            const double d = valid ? v[c] - smean[c] : 0.0;
            const double q = wave_sum(d * d);
            if (lane == 0) red[c][wv][0] = q;
        }
    }
    __syncthreads();
    if (threadIdx.x < nc) {
        const int c = threadIdx.x;
        f->mpq[(size_t)(o0 + c) * T + tile] = (red[c][0][0] + red[c][1][0]) + (red[c][2][0] + red[c][3][0]);
    }
    if (blockIdx.y == 0 && threadIdx.x == 0) f->mpw[tile] = sw;
}

/* pairwise tree (contiguous, zero padded to a power of two) over the T tile values of a
 * filter, in chunks of 256 then over the chunks; every thread gets the total */
__device__ double tile_tree(const double *v, int T, double *lds, double *chunks)
{
    const int nch = (T + TILE - 1) / TILE;
    for (int c = 0; c < nch; c++) {
        const int t = c * TILE + threadIdx.x;
        const double s = tree256(t < T ? v[t] : 0.0, lds);
        if (threadIdx.x == 0) chunks[c] = s;
    }
    __syncthreads();
    __shared__ double r;
    if (threadIdx.x == 0) {
        int P = 1;
        while (P < nch) P <<= 1;
        for (int c = nch; c < P; c++) chunks[c] = 0.0;
        for (int h = 1; h < P; h <<= 1)
            for (int i = 0; i < P; i += 2 * h) chunks[i] = chunks[i] + chunks[i + h];
        r = chunks[0];
    }
    __syncthreads();
    const double out = r;
    __syncthreads();
    return out;
}

/* one block per (filter, component): weighted mean sum(w theta)/sum(w) and variance
 * M2/(J-1), the (count, mean, M2) of the tiles merged along the same fixed tree */
__global__ __launch_bounds__(TILE) void k_mif_reduce(plom_dfilter *filters, const int *fidx, int n)
{
    __shared__ double lds[8];
    __shared__ double chunks[MAX_TILES / TILE + 1];
    __shared__ double sc[TILE], sm[TILE], sq[TILE];
    __shared__ double cc[MAX_TILES / TILE + 1], cm[MAX_TILES / TILE + 1], cq[MAX_TILES / TILE + 1];
    const int fi = fidx ? fidx[blockIdx.x] : (int)blockIdx.x;
    plom_dfilter *f = &filters[fi];
    if (n >= f->N || !(f->flags & PLOM_FLAG_MIF) || f->nonan[n] == 0) return;
    const int J = f->J, T = f->T, nth = f->n_theta;
    const int o = blockIdx.y;
    if (o >= nth) return;
    const double W = tile_tree(f->mpw, T, lds, chunks);
    const int nch = (T + TILE - 1) / TILE;
    {
        const double SW = tile_tree(f->mps + (size_t)o * T, T, lds, chunks);
        for (int c = 0; c < nch; c++) {
            const int t = c * TILE + threadIdx.x;
            if (t < T) {
                const int rem = J - t * TILE;
                sc[threadIdx.x] = (double)(rem < TILE ? rem : TILE);
                sm[threadIdx.x] = f->mpm[(size_t)o * T + t];
                sq[threadIdx.x] = f->mpq[(size_t)o * T + t];
            } else {
                sc[threadIdx.x] = 0.0; sm[threadIdx.x] = 0.0; sq[threadIdx.x] = 0.0;
            }
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
            f->mmean[(size_t)(n + 1) * nth + o] = SW / W;
            f->mvar[(size_t)(n + 1) * nth + o] = cq[0] / (double)(J - 1);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(TILE) void k_scan(plom_dfilter *filters, const int2 *tiles)
{
    __shared__ uint64_t lds[8];
    int fi = 0, tile = blockIdx.x;
    if (tiles) { const int2 tt = tiles[blockIdx.x]; fi = tt.x; tile = tt.y; }
    plom_dfilter *f = &filters[fi];
    const int n = f->step;
    if (n >= f->N) return;
    const bool active = !(f->flags & PLOM_FLAG_NO_FILTER) && f->nonan[n] > 0 && f->rflag[n];
    if (!active) return;
    const double m = f->scal[0], S = f->scal[1];
    const int j = tile * TILE + threadIdx.x;
    const double l = f->logw[j];
    uint64_t q = 0;
    double e = 0.0;
    if (l != -PLOM_INF) {
        e = plom_exp(l - m);
        q = (uint64_t)((e / S) * 0x1p62);
    }
    if (f->mw) f->mw[j] = e;       /* PLOM_FLAG_MIF: the weights of the theta moments */
    const uint64_t c = scan256(q, lds);
    f->cum[j] = c;
    if (threadIdx.x == TILE - 1) f->ttot[tile] = c;
}

/* one block per filter: inclusive prefix of the tile totals, then the next interval */
__global__ __launch_bounds__(TILE) void k_advance(plom_dfilter *filters, const int *fidx)
{
    __shared__ uint64_t lds[8];
    __shared__ uint64_t s_carry;
    const int fi = fidx ? fidx[blockIdx.x] : (int)blockIdx.x;
    plom_dfilter *f = &filters[fi];
    const int n = f->step;
    if (n >= f->N) return;
    const bool active = !(f->flags & PLOM_FLAG_NO_FILTER) && f->nonan[n] > 0 && f->rflag[n];
    if (active) {
        if (threadIdx.x == 0) s_carry = 0;
        __syncthreads();
        for (int base = 0; base < f->T; base += TILE) {
            const int t = base + threadIdx.x;
            const uint64_t v = t < f->T ? f->ttot[t] : 0ull;
            const uint64_t c = scan256(v, lds) + s_carry;
            if (t < f->T) f->tincl[t] = c;
            __syncthreads();
            if (threadIdx.x == TILE - 1) s_carry = c;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) f->step = n + 1;
}

/* resampling of the last interval: ancestry row N-1 and the resampled final cloud */
template <typename R>
__global__ __launch_bounds__(TILE) void k_final(plom_dfilter *filters, const int2 *tiles, int S)
{
    int fi = 0, tile = blockIdx.x;
    if (tiles) { const int2 tt = tiles[blockIdx.x]; fi = tt.x; tile = tt.y; }
    plom_dfilter *f = &filters[fi];
    const int N = f->N, J = f->J, Jp = f->Jp;
    const int j = tile * TILE + threadIdx.x;
    if (j >= J || N == 0) return;
    int a = j;
    if (f->rflag[N - 1]) a = select_ancestor(f, j, N - 1);
    if (f->ancestry) f->ancestry[(size_t)(N - 1) * J + j] = a;
    const R *Xin = (const R *)f->X[N & 1];
    R *Xout = (R *)f->X[(N + 1) & 1];
    for (int k = 0; k < S; k++) Xout[(size_t)k * Jp + j] = Xin[(size_t)k * Jp + a];
}


/* block k: the weighted mean of quantity k (a tree over the tiles' tree sums) */
__global__ __launch_bounds__(TILE) void k_hat_mean(plom_dfilter *f)
{
    __shared__ double lds[8];
    __shared__ double s_chunk[MAX_TILES / TILE + 1];
    const int n = f->step - 1;
    if (n < 0 || n >= f->N || !f->hkey) return;
    const int k = blockIdx.x;
    const int T = f->T, NH = f->n_hat;
    const uint64_t Q = hat_weighted(f, n) ? f->tincl[T - 1] : (uint64_t)f->J;
    double *out = f->hat + ((size_t)(n + 1) * NH + k) * 3;
    {
        const int nch = (T + TILE - 1) / TILE;
        for (int c = 0; c < nch; c++) {
            const int t = c * TILE + threadIdx.x;
            const double v = tree256(t < T ? f->hpart[(size_t)k * T + t] : 0.0, lds);
            if (threadIdx.x == 0) s_chunk[c] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int P = 1;
            while (P < nch) P <<= 1;
            for (int c = nch; c < P; c++) s_chunk[c] = 0.0;
            for (int h = 1; h < P; h <<= 1)
                for (int i = 0; i < P; i += 2 * h) s_chunk[i] = s_chunk[i] + s_chunk[i + h];
            out[1] = s_chunk[0] / (double)Q;
        }
    }
}

/* exact sum of a uint64 over a block of HAT_BOUND_THREADS (every thread gets it) */
#define HAT_BOUND_THREADS 1024
__device__ inline uint64_t sum_block_u64(uint64_t v, uint64_t *lds)
{
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) v += __shfl_xor(v, h, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    uint64_t r = 0;
    for (int i = 0; i < HAT_BOUND_THREADS / 64; i++) r += lds[i];
    __syncthreads();
    return r;
}

/* block-wide minimum of lo and maximum of hi (lds: 2 HAT_BOUND_THREADS / 64 entries) */
__device__ inline void minmax_block_u64(uint64_t *lo, uint64_t *hi, uint64_t *lds)
{
    uint64_t a = *lo, b = *hi;
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) {
        const uint64_t a2 = __shfl_xor(a, h, 64), b2 = __shfl_xor(b, h, 64);
        a = a2 < a ? a2 : a;
        b = b2 > b ? b2 : b;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { lds[2 * w] = a; lds[2 * w + 1] = b; }
    __syncthreads();
    a = ~0ull;
    b = 0ull;
    for (int i = 0; i < HAT_BOUND_THREADS / 64; i++) {
        a = lds[2 * i] < a ? lds[2 * i] : a;
        b = lds[2 * i + 1] > b ? lds[2 * i + 1] : b;
    }
    __syncthreads();
    *lo = a;
    *hi = b;
}

/* blocks 2k, 2k+1: low95 and high95 of quantity k, the smallest key whose weight prefix
 * reaches 1/40 (39/40) of the total, by bisection of the 64-bit key space; in every round
 * each thread counts, in the tiles it covers, the weight of the keys at most the midpoint
 * (a binary search of the sorted tile), and the block sums the counts (exact integers, so
 * the order does not matter).  HAT_BOUND_THREADS threads, so that a thread covers one
 * tile per round up to 1024 tiles. */
__global__ __launch_bounds__(HAT_BOUND_THREADS) void k_hat_bounds(plom_dfilter *f)
{
    __shared__ uint64_t lds64[2 * (HAT_BOUND_THREADS / 64)];
    const int n = f->step - 1;
    if (n < 0 || n >= f->N || !f->hkey) return;
    const int k = blockIdx.x / 2, role = 1 + (int)(blockIdx.x % 2);
    const int T = f->T, Jp = f->Jp, NH = f->n_hat;
    const uint64_t Q = hat_weighted(f, n) ? f->tincl[T - 1] : (uint64_t)f->J;
    double *out = f->hat + ((size_t)(n + 1) * NH + k) * 3;
    const uint64_t pnum = role == 1 ? 1u : 39u;
    const uint64_t *keys = f->hkey + (size_t)k * Jp, *pref = f->hpref + (size_t)k * Jp;
    /* The bound lies between the smallest and the largest of the tiles' own bounds (the
     * smallest key of a tile whose prefix reaches pnum/40 of the tile's weight), over the
     * tiles with weight: below the smallest every such tile falls short of its share, so
     * the sum does; at the largest every tile reaches its share, so the sum does.  The
     * bisection starts from that bracket (the full key range when no tile has weight), and
     * finds the same key. */
    uint64_t tlo = ~0ull, thi = 0ull;
    for (int t = threadIdx.x; t < T; t += HAT_BOUND_THREADS) {
        const uint64_t *pt = pref + (size_t)t * TILE;
        const uint64_t Wt = pt[TILE - 1];
        if (!Wt) continue;
        int a = 0, b = TILE - 1;                 /* first position whose prefix reaches the share */
        while (a < b) {
            const int m = (a + b) >> 1;
            if ((plom_u128)pt[m] * 40u >= (plom_u128)Wt * pnum) b = m; else a = m + 1;
        }
        const uint64_t x = keys[(size_t)t * TILE + a];
        tlo = x < tlo ? x : tlo;
        thi = x > thi ? x : thi;
    }
    minmax_block_u64(&tlo, &thi, lds64);
    uint64_t lo = 0, hi = ~0ull;
    if (tlo <= thi) { lo = tlo; hi = thi; }
    /* the last key of each 16-key segment of the thread's first tile, in registers: the
     * segments wholly at most the midpoint are counted there, and a round's search goes
     * to memory only inside the one segment that straddles it */
    uint64_t seg[TILE / 16];
    const bool own = (int)threadIdx.x < T;
#pragma unroll
    for (int i = 0; i < TILE / 16; i++) seg[i] = own ? keys[(size_t)threadIdx.x * TILE + 16 * i + 15] : 0ull;
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        uint64_t W = 0;
        for (int t = threadIdx.x; t < T; t += HAT_BOUND_THREADS) {
            const uint64_t *kt = keys + (size_t)t * TILE;
            int a = 0, b = TILE;                 /* count of keys <= mid in the sorted tile */
            if (t == (int)threadIdx.x) {
                int ns = 0;                      /* segments whose keys are all <= mid */
#pragma unroll
                for (int i = 0; i < TILE / 16; i++) ns += seg[i] <= mid ? 1 : 0;
                a = 16 * ns;                     /* kt[a - 1] <= mid, and kt[a + 15] > mid */
                b = ns < TILE / 16 ? a + 15 : TILE;
            }
            while (a < b) {
                const int m = (a + b) >> 1;
                if (kt[m] <= mid) a = m + 1; else b = m;
            }
            if (a) W += pref[(size_t)t * TILE + a - 1];
        }
        W = sum_block_u64(W, lds64);
        if ((plom_u128)W * 40u >= (plom_u128)Q * pnum) hi = mid; else lo = mid + 1;
    }
    if (threadIdx.x == 0) out[role == 1 ? 0 : 2] = hat_unkey(lo);
}


__global__ void k_selftest_math(int n, const double *in, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = in[i];
    out[i * 8 + 0] = plom_exp(x);
    out[i * 8 + 1] = plom_log(x);
    out[i * 8 + 2] = plom_Phi(x);
    out[i * 8 + 3] = plom_sin(x);
    out[i * 8 + 4] = plom_cos(x);
    out[i * 8 + 5] = plom_sqrt(x);
    out[i * 8 + 6] = plom_sqrt_mid(x);
    out[i * 8 + 7] = plom_sqrt_rate(x);
}

/* plain copy (16-byte vectors, grid-stride): the bandwidth reference and the calibration
 * of the FETCH_SIZE / WRITE_SIZE counters */
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

/* the first n Gaussians of a particle's stream over interval n0, FP64 (fp32 = 0) or FP32
 * (fp32 = 1, widened to double) */
__global__ void k_selftest_normals(uint32_t seed, uint32_t stream, uint32_t n0, uint32_t j, int n, int fp32, double *out)
{
    __shared__ float s_tab32[PLOM_ICDF_TAB_SIZE];
    for (int i = threadIdx.x; i < PLOM_ICDF_TAB_SIZE; i += blockDim.x) s_tab32[i] = plom_icdf_tab32[i];
    __syncthreads();
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (fp32) {
        plom_nstream32 ns;
        plom_nstream32_init(&ns, seed, stream, n0, j, s_tab32);
        for (int i = 0; i < n; i += 2) {
            float a, b;
            plom_normal_pair32(&ns, &a, &b);
            out[i] = a;
            if (i + 1 < n) out[i + 1] = b;
        }
        return;
    }
    plom_nstream64 ns;
    plom_nstream64_init(&ns, seed, stream, n0, j, plom_icdf_tab);
    for (int i = 0; i + 1 < n; i += 2) plom_normal_pair64(&ns, &out[i], &out[i + 1]);
    if (n & 1) {
        double spare;
        plom_normal_pair64(&ns, &out[n - 1], &spare);
    }
}


/* ------------------------------------------------------------------ model table */
/* the Box-Muller table for host code (the kernels read the __constant__ one) */
static const double g_bm_tab_host[PLOM_BM_TAB_SIZE] = PLOM_BM_TAB_VALUES;

/* the models compiled into the library (one translation unit each), then the models
 * registered by plom_model_load, in id order; entries are never removed, so a reader
 * needs only the published count */
#define PLOM_MAX_MODELS 64
#define DECLARE_BUILTIN(lname) extern "C" const model_entry *plom_model_entry_##lname(void);
PLOM_BUILTIN_MODELS(DECLARE_BUILTIN)
#define LIST_BUILTIN(lname) plom_model_entry_##lname,
static const model_entry *(*const g_builtin[])(void) = {PLOM_BUILTIN_MODELS(LIST_BUILTIN)};
static const model_entry *g_model_tab[PLOM_MAX_MODELS];
static std::atomic<int> g_model_count{0};
static std::mutex g_model_mu;

static int model_count()
{
    static std::once_flag once;
    std::call_once(once, [] {
        int n = 0;
        for (auto fn : g_builtin) g_model_tab[n++] = fn();
        g_model_count.store(n, std::memory_order_release);
    });
    return g_model_count.load(std::memory_order_acquire);
}

/* model `id`, or null when out of range */
static const model_entry *model_at(int id)
{
    return (id >= 0 && id < model_count()) ? g_model_tab[id] : nullptr;
}

extern "C" int plom_model_count(void) { return model_count(); }

extern "C" int plom_model_find(const char *name)
{
    const int n = model_count();
    for (int i = 0; i < n; i++)
        if (name && !strcmp(name, g_model_tab[i]->name)) return i;
    return set_err(PLOM_ERR_MODEL, "unknown model '%s'", name ? name : "(null)");
}

extern "C" int plom_model_name(int id, char *buf, int len)
{
    const model_entry *me = model_at(id);
    if (!me) return set_err(PLOM_ERR_MODEL, "model id %d out of range", id);
    if (buf && len > 0) snprintf(buf, (size_t)len, "%s", me->name);
    return PLOM_OK;
}

extern "C" int plom_model_info_get(int id, plom_model_info *out)
{
    const model_entry *me = model_at(id);
    if (!me) return set_err(PLOM_ERR_MODEL, "model id %d out of range", id);
    if (!out) return set_err(PLOM_ERR_ARG, "null output");
    *out = me->info;
    return PLOM_OK;
}

/* Register the model of a model library built by plom_build.py (plom_model.hip compiled
 * with PLOM_MODEL_PLUGIN): dlopen it, check that it was compiled against this library's
 * device descriptor, and append it to the table.  Loading a library whose model name is
 * already registered gives that model's id again. */
extern "C" int plom_model_load(const char *path, int32_t *id)
{
    if (!path || !id) return set_err(PLOM_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lock(g_model_mu);
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return set_err(PLOM_ERR_MODEL, "cannot load model library %s: %s", path, dlerror());
    typedef const model_entry *(*entry_fn)(void);
    entry_fn fn = (entry_fn)dlsym(h, "plom_model_plugin");
    if (!fn) {
        dlclose(h);
        return set_err(PLOM_ERR_MODEL, "%s exports no plom_model_plugin (not a model library)", path);
    }
    const model_entry *me = fn();
    if (!me || me->abi != PLOM_MODEL_ABI || me->dfilter_bytes != (int32_t)sizeof(plom_dfilter)) {
        dlclose(h);
        return set_err(PLOM_ERR_MODEL, "%s was compiled against another version of the library (abi %d, descriptor "
                       "%d bytes; this library: abi %d, %d bytes)", path, me ? me->abi : -1, me ? me->dfilter_bytes : -1,
                       PLOM_MODEL_ABI, (int)sizeof(plom_dfilter));
    }
    const int n = model_count();
    for (int i = 0; i < n; i++)
        if (!strcmp(g_model_tab[i]->name, me->name)) {
            /* the same model again: keep the first registration (its library stays open) */
            if (g_model_tab[i] != me) dlclose(h);
            *id = i;
            return PLOM_OK;
        }
    if (n >= PLOM_MAX_MODELS) {
        dlclose(h);
        return set_err(PLOM_ERR_MODEL, "at most %d models", PLOM_MAX_MODELS);
    }
    g_model_tab[n] = me;   /* the library stays open until the process ends */
    g_model_count.store(n + 1, std::memory_order_release);
    *id = n;
    return PLOM_OK;
}

/* the hat of one interval: the model's tile sort and residuals, then the selection */
static void launch_hat_all(const model_entry *me, int prec, hipStream_t s, plom_dfilter *df, int ntiles, int n_hat)
{
    me->launch_hat(prec, s, df, ntiles, n_hat);
    hipLaunchKernelGGL(k_hat_mean, dim3(n_hat), dim3(TILE), 0, s, df);
    hipLaunchKernelGGL(k_hat_bounds, dim3(2 * n_hat), dim3(HAT_BOUND_THREADS), 0, s, df);
}

/* ------------------------------------------------------------------ context / filters */
struct plom_ctx {
    int device;
    hipStream_t stream;
    hipStream_t io;                 /* result reads that must not wait for queued work */
    hipEvent_t ev[4];
    double last_total_ms, last_prop_ms, last_scan_ms;
    int ktiming;                    /* bracket every launch with events */
    std::vector<hipEvent_t> kev;    /* pool: 3 per launch round and lane */
    struct kspan { int kind, lane, a, b; };   /* kind 0: propagate launches, 1: weights / scan */
    std::vector<kspan> kspans;      /* event pairs of the last run (timing on) */
    int kspan_rounds;               /* launch rounds they cover */
    int kprop_launches;             /* propagate kernel launches of the last run */
    std::vector<hipStream_t> lane;  /* batch lane streams, created on first use ([0] unused: lane 0 is `stream`) */
    std::vector<hipEvent_t> join_ev;
    hipEvent_t fork_ev;
    hipStream_t mside;              /* the theta moments of MIF filters beside the next propagate */
    hipEvent_t mside_fork, mside_done;
    int tpb;                        /* tiles per propagate block of single filters, the server and
                                       batch groups without a tuned value */
    char *scratch;                  /* this device's fixed scratch (plom_smc_run_noalloc), or null */
};

/* event k of the per-launch timing pool (grown on demand), recorded on stream s */
static int kev_record(plom_ctx *c, int k, hipStream_t s)
{
    while ((int)c->kev.size() <= k) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->kev.push_back(e);
    }
    HIPCHK(hipEventRecord(c->kev[(size_t)k], s));
    return PLOM_OK;
}

/* timing on: an event before the launches of a span; returns its pool index (or -1) */
static int kspan_open(plom_ctx *c, int *k, hipStream_t s)
{
    if (!c->ktiming) return -1;
    if (kev_record(c, *k, s) != PLOM_OK) return -2;
    return (*k)++;
}

/* timing on: an event after the launches of the span opened at `a` */
static int kspan_close(plom_ctx *c, int *k, hipStream_t s, int a, int kind, int lane)
{
    if (!c->ktiming) return -1;
    if (a < 0 || kev_record(c, *k, s) != PLOM_OK) return -2;
    c->kspans.push_back(plom_ctx::kspan{kind, lane, a, *k});
    return (*k)++;
}

struct plom_filter {
    plom_filter_desc desc;
    plom_model_info info;
    int variant;
    char *dev;              /* one allocation holding everything (or the fixed scratch) */
    int fixed_dev;          /* dev is the device's fixed scratch: nothing to free */
    size_t dev_bytes;
    int async_alloc;        /* allocated with hipMallocAsync (server) */
    plom_dfilter *df;       /* device copy of the descriptor */
    plom_dfilter h;         /* host copy (device pointers) */
    int has_run;
    int dirty;              /* descriptor changed on the device since batch plans copied it */
    size_t off_pv, off_pvf, off_x0;   /* image offsets of the parameter and initial-state inputs */
    size_t off_mutfac;      /* PLOM_FLAG_MIF: image offset of mut_fac */
    int mif_ready;          /* PLOM_FLAG_MIF: a starting cloud is uploaded and not yet consumed */
    std::vector<int32_t> nonan;       /* per interval: observed series (host copy) */
    std::vector<double> pv_h, times_h, cov_h;   /* host copies for the stored clouds' observations */
    int cov_kmax_h;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

/* The segments of the lineage replay: K intervals each (at least 8, and at most
 * PLOM_LIN_SEG_MAX segments), ceil(N / K) of them; the cloud is copied after intervals
 * K - 1, 2K - 1, ... for every segment after the first.  The ode implementation replays
 * its lineage in one segment. */
static void lineage_segments(int N, uint32_t flags, int *K, int *nseg)
{
    if (N <= 0) { *K = 1; *nseg = 0; return; }
    if (flags & PLOM_FLAG_ODE) { *K = N; *nseg = 1; return; }
    int k = (N + PLOM_LIN_SEG_MAX - 1) / PLOM_LIN_SEG_MAX;
    if (k < 8) k = 8;
    *K = k;
    *nseg = (N + k - 1) / k;
}

/* One device allocation per filter.  The work buffers come first; the tail from off_img
 * on (descriptor, counters, per-interval outputs, inputs) is written from one host image
 * in a single copy, so that creating a filter is one upload (from pinned memory when the
 * server creates it) and starts from zeroed counters without memset calls. */
struct layout {
    size_t off_X0, off_X1, off_logw, off_cum, off_tmax, off_tsum, off_tsum2, off_ttot, off_tincl, off_anc,
        off_lslots, off_lstate, off_lobs, off_lsnap, off_hkey, off_hpref, off_hpart, off_hat, off_pres, off_prw, off_prp,
        off_th0, off_th1, off_mw, off_mpw, off_mps, off_mpm, off_mpq, off_mmean, off_mvar,
        off_img, off_df, off_cnt, off_scal, off_naf, off_llike_n, off_ess_n, off_rflag, off_nonan, off_times,
        off_data, off_pv, off_pvf, off_x0, off_pm_src, off_pm_tr, off_pm_sc, off_pm_dur, off_pm_min, off_pm_max,
        off_pm_mult, off_pm_sd, off_mutfac, off_cov, off_dm_group, off_dm_ng, off_dm_tr, off_dm_sc, off_dm_dur,
        off_dm_min, off_dm_max, off_dm_mult, off_dm_vol, off_allX, off_pm_chol, off_pm_prior, off_pm_ppow,
        off_pm_puntil, off_spl_n, off_spl_off, off_spl_val, total;   /* off_dm_vol: int32 vol slots */
};

/* cov_vals: values of the covariate table (0 without time-varying covariates); chol_vals: of
 * the MIF walk's covariance factor (0 without one); prior_on: the MIF prior patch; spl_vals:
 * values of the covariate splines (the ode implementation with time-varying covariates) */
static int make_layout(const plom_filter_desc *d, const plom_model_info *mi, layout *L, size_t cov_vals = 0,
                       size_t chol_vals = 0, int prior_on = 0, size_t spl_vals = 0)
{
    const size_t J = (size_t)d->J, N = (size_t)d->N;
    const size_t Jp = (J + TILE - 1) / TILE * TILE, T = Jp / TILE;
    const size_t rs = d->precision == PLOM_FP32 ? 4 : 8;
    const size_t S = (size_t)mi->size_proj;
    size_t o = 0;
#define PUT(field, bytes) do { L->field = o; o += align256(bytes); } while (0)
    PUT(off_X0, S * Jp * rs);
    PUT(off_X1, S * Jp * rs);
    PUT(off_logw, Jp * 8);
    PUT(off_cum, Jp * 8);
    PUT(off_tmax, T * 8);
    PUT(off_tsum, T * 8);
    PUT(off_tsum2, T * 8);
    PUT(off_ttot, T * 8);
    PUT(off_tincl, T * 8);
    PUT(off_anc, (d->flags & (PLOM_FLAG_STORE_ANCESTRY | PLOM_FLAG_LINEAGE | PLOM_FLAG_ALL_STATES)) ? N * J * 4 : 0);
    PUT(off_allX, (d->flags & PLOM_FLAG_ALL_STATES) ? (N + 1) * S * Jp * rs : 0);
    PUT(off_lslots, (d->flags & PLOM_FLAG_LINEAGE) ? N * 4 : 0);
    PUT(off_lstate, (d->flags & PLOM_FLAG_LINEAGE) ? (N + 1) * S * 8 : 0);
    PUT(off_lobs, (d->flags & PLOM_FLAG_LINEAGE) ? (N + 1) * (size_t)mi->n_ts * 2 * 8 : 0);
    {
        int K, nseg;
        lineage_segments((int)N, d->flags, &K, &nseg);
        PUT(off_lsnap, ((d->flags & PLOM_FLAG_LINEAGE) && nseg > 1) ? (size_t)(nseg - 1) * S * Jp * rs : 0);
    }
    const size_t NH = (d->flags & PLOM_FLAG_HAT) ? (size_t)mi->n_hat : 0;
    PUT(off_hkey, NH * Jp * 8);
    PUT(off_hpref, NH * Jp * 8);
    PUT(off_hpart, NH * T * 8);
    PUT(off_hat, NH * (N + 1) * 3 * 8);
    PUT(off_pres, NH ? N * (size_t)mi->n_ts * 2 * 8 : 0);
    PUT(off_prw, NH ? T * 8 : 0);
    PUT(off_prp, NH ? 4 * (size_t)mi->n_ts * T * 8 : 0);
    const size_t NT = (d->flags & PLOM_FLAG_MIF) ? (size_t)d->n_theta : 0;
    PUT(off_th0, NT * Jp * 8);
    PUT(off_th1, NT * Jp * 8);
    PUT(off_mw, NT ? Jp * 8 : 0);
    PUT(off_mpw, NT ? T * 8 : 0);
    PUT(off_mps, NT * T * 8);
    PUT(off_mpm, NT * T * 8);
    PUT(off_mpq, NT * T * 8);
    PUT(off_mmean, NT * (N + 1) * 8);
    PUT(off_mvar, NT * (N + 1) * 8);
    L->off_img = o;
    PUT(off_df, sizeof(plom_dfilter));
    PUT(off_cnt, 2 * 4);
    PUT(off_scal, 4 * 8);
    PUT(off_naf, 4);
    PUT(off_llike_n, N * 8);
    PUT(off_ess_n, N * 8);
    PUT(off_rflag, N * 4);
    PUT(off_nonan, N * 4);
    PUT(off_times, (N + 1) * 8);
    PUT(off_data, N * (size_t)mi->n_ts * 8);
    PUT(off_pv, (size_t)mi->pv_size * 8);
    PUT(off_pvf, (size_t)mi->pv_size * 4);
    PUT(off_x0, S * 8);
    const size_t NPP = NT ? (size_t)mi->n_pp : 0;
    PUT(off_pm_src, NPP * 4);
    PUT(off_pm_tr, NT * 4);
    PUT(off_pm_sc, NT * 4);
    PUT(off_pm_dur, NT * 4);
    PUT(off_pm_min, NT * 8);
    PUT(off_pm_max, NT * 8);
    PUT(off_pm_mult, NT * 8);
    PUT(off_pm_sd, NT * 8);
    PUT(off_mutfac, NT ? N * 8 : 0);
    PUT(off_pm_chol, chol_vals * 8);
    PUT(off_pm_prior, prior_on ? NT * 4 : 0);
    PUT(off_pm_ppow, prior_on ? NT * 8 : 0);
    PUT(off_pm_puntil, prior_on ? NT * 4 : 0);
    PUT(off_spl_n, spl_vals ? (size_t)mi->n_fix * 4 : 0);
    PUT(off_spl_off, spl_vals ? (size_t)mi->n_fix * 4 : 0);
    PUT(off_spl_val, spl_vals * 8);
    PUT(off_cov, cov_vals * 8);
    const size_t ND = (size_t)mi->n_drift * (size_t)mi->n_cac;
    PUT(off_dm_group, ND * 4);
    PUT(off_dm_ng, (size_t)mi->n_drift * 4);
    PUT(off_dm_tr, ND * 4);
    PUT(off_dm_sc, ND * 4);
    PUT(off_dm_dur, ND * 4);
    PUT(off_dm_min, ND * 8);
    PUT(off_dm_max, ND * 8);
    PUT(off_dm_mult, ND * 8);
    PUT(off_dm_vol, ND * 4);
#undef PUT
    L->total = o;
    return PLOM_OK;
}

static int check_desc(const plom_filter_desc *d, plom_model_info *mi)
{
    if (!d) return set_err(PLOM_ERR_ARG, "null descriptor");
    const model_entry *me = model_at(d->model_id);
    if (!me) return set_err(PLOM_ERR_MODEL, "model id %d out of range", d->model_id);
    *mi = me->info;
    if (d->J < 1 || d->J > PLOM_MAX_J) return set_err(PLOM_ERR_ARG, "J=%d outside [1, %d]", d->J, PLOM_MAX_J);
    if (d->N < 0) return set_err(PLOM_ERR_ARG, "N=%d < 0", d->N);
    if (d->precision != PLOM_FP64 && d->precision != PLOM_FP32) return set_err(PLOM_ERR_ARG, "bad precision %d", d->precision);
    if (!(d->dt > 0.0)) return set_err(PLOM_ERR_ARG, "dt must be > 0");
    if (!(d->like_min > 0.0)) return set_err(PLOM_ERR_ARG, "like_min must be > 0");
    if ((d->flags & PLOM_FLAG_PSR) && d->precision != PLOM_FP64)
        return set_err(PLOM_ERR_ARG, "PLOM_FLAG_PSR (the psr implementation) runs in FP64 only");
    if (d->flags & PLOM_FLAG_ODE) {
        if (d->precision != PLOM_FP64) return set_err(PLOM_ERR_ARG, "PLOM_FLAG_ODE (the ode implementation) runs in FP64 only");
        if (d->flags & PLOM_FLAG_PSR) return set_err(PLOM_ERR_ARG, "PLOM_FLAG_ODE and PLOM_FLAG_PSR exclude each other");
        if (d->eps_abs < 0.0 || d->eps_rel < 0.0 || !(d->eps_abs + d->eps_rel >= 0.0))
            return set_err(PLOM_ERR_ARG, "eps_abs and eps_rel must be >= 0");
    }
    if (d->flags & PLOM_FLAG_MIF) {
        if (d->n_theta < 1 || d->n_theta > 4096) return set_err(PLOM_ERR_ARG, "n_theta=%d outside [1, 4096]", d->n_theta);
        if (d->flags & (PLOM_FLAG_NO_FILTER | PLOM_FLAG_HAT | PLOM_FLAG_LINEAGE))
            return set_err(PLOM_ERR_ARG, "PLOM_FLAG_MIF does not combine with NO_FILTER, HAT or LINEAGE");
    }
    return PLOM_OK;
}

extern "C" int plom_covariate_eval(const plom_covariate *cov, int nt, const double *t, double *out)
{
    if (!cov || nt < 0 || (nt > 0 && (!t || !out))) return set_err(PLOM_ERR_ARG, "null argument");
    char msg[256];
    if (plom_cov_check(cov, 1, msg, sizeof msg)) return set_err(PLOM_ERR_ARG, "%s", msg);
    std::vector<double> c((size_t)cov->n + 1);
    if (plom_spline_prepare(cov->n, cov->x, cov->y, c.data())) return set_err(PLOM_ERR_NOMEM, "host allocation failed");
    for (int i = 0; i < nt; i++) {
        if (cov->n > 1 && !(t[i] >= cov->x[0] && t[i] <= cov->x[cov->n - 1]))
            return set_err(PLOM_ERR_ARG, "time %.17g outside the covariate's knots [%.17g, %.17g]", t[i], cov->x[0],
                           cov->x[cov->n - 1]);
        out[i] = plom_spline_eval(cov->n, cov->x, cov->y, c.data(), t[i]);
    }
    return PLOM_OK;
}

extern "C" int64_t plom_filter_device_bytes(const plom_filter_desc *d)
{
    plom_model_info mi;
    if (check_desc(d, &mi) != PLOM_OK) return -1;
    layout L;
    make_layout(d, &mi, &L);
    return (int64_t)L.total;
}

/* ---------------------------------------------------------------- fixed device scratch */
/* The device memory of plom_smc_run_noalloc: a file-scope array that every GPU the library
 * runs on holds from the library's load until the process ends.  In host code the name
 * `scratch` is only a host-side stand-in for the symbol, not a device address: the address
 * on each GPU comes from hipGetSymbolAddress, taken once per device with that device current
 * (in plom_ctx_create) and kept in g_scratch_dev; host code uses only that pointer.  Calls
 * share their GPU's array and find the previous call's contents in it. */
static __device__ __align__(256) char scratch[PLOM_NOALLOC_MAX_BYTES];
#define PLOM_SCRATCH_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static char *g_scratch_dev[PLOM_SCRATCH_MAX_DEVICES];

/* the scratch's address on `device` (the current device), looked up on first use */
static int scratch_bind(int device, char **out)
{
    if (device < 0 || device >= PLOM_SCRATCH_MAX_DEVICES)
        return set_err(PLOM_ERR_ARG, "device %d: the fixed scratch serves devices 0..%d", device, PLOM_SCRATCH_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_err(PLOM_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch_dev[device] = (char *)ptr;
    }
    *out = g_scratch_dev[device];
    return PLOM_OK;
}

extern "C" int plom_ctx_create(int device, plom_ctx **out)
{
    if (!out) return set_err(PLOM_ERR_ARG, "null output");
    HIPCHK(hipSetDevice(device));
    char *scratch_dev = nullptr;
    if (device < PLOM_SCRATCH_MAX_DEVICES) {
        const int rc = scratch_bind(device, &scratch_dev);
        if (rc != PLOM_OK) return rc;
    }
    void *mem = calloc(1, sizeof(plom_ctx));
    if (!mem) return set_err(PLOM_ERR_NOMEM, "host allocation failed");
    plom_ctx *c = new (mem) plom_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { free(c); return set_err(PLOM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    e = hipStreamCreateWithFlags(&c->io, hipStreamNonBlocking);
    if (e != hipSuccess) { (void)hipStreamDestroy(c->stream); free(c); return set_err(PLOM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    for (int i = 0; i < 4; i++) HIPCHK(hipEventCreate(&c->ev[i]));
    c->ktiming = 0;
    c->kspan_rounds = 0;
    c->tpb = 1;
    c->scratch = scratch_dev;
    *out = c;
    return PLOM_OK;
}

extern "C" int plom_ctx_destroy(plom_ctx *c)
{
    if (!c) return PLOM_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 4; i++) HIPCHK(hipEventDestroy(c->ev[i]));
    for (auto e : c->kev) HIPCHK(hipEventDestroy(e));
    for (size_t l = 0; l < c->lane.size(); l++) {
        HIPCHK(hipStreamSynchronize(c->lane[l]));
        HIPCHK(hipStreamDestroy(c->lane[l]));
        HIPCHK(hipEventDestroy(c->join_ev[l]));
    }
    if (c->fork_ev) HIPCHK(hipEventDestroy(c->fork_ev));
    if (c->mside) {
        HIPCHK(hipStreamSynchronize(c->mside));
        HIPCHK(hipStreamDestroy(c->mside));
        HIPCHK(hipEventDestroy(c->mside_fork));
        HIPCHK(hipEventDestroy(c->mside_done));
    }
    HIPCHK(hipStreamDestroy(c->stream));
    HIPCHK(hipStreamDestroy(c->io));
    c->~plom_ctx();
    free(c);
    return PLOM_OK;
}

extern "C" int plom_sync(plom_ctx *c)
{
    if (!c) return set_err(PLOM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return PLOM_OK;
}

/* back to interval 0 for every filter of a batch (a kernel, so that it can sit inside a
 * captured graph: memset nodes are not used in graphs) */
__global__ void k_reset(plom_dfilter *filters, int nf)
{
    const int fi = blockIdx.x;
    if (fi >= nf) return;
    plom_dfilter *f = &filters[fi];
    for (int n = threadIdx.x; n < f->N; n += blockDim.x) {
        f->rflag[n] = 0;
        f->llike_n[n] = 0.0;
        f->ess_n[n] = 0.0;
    }
    if (threadIdx.x == 0) {
        f->cnt[0] = 0u;
        f->cnt[1] = 0u;
        for (int k = 0; k < 4; k++) f->scal[k] = 0.0;
        f->n_all_fail[0] = 0;
        f->step = 0;
    }
}

/* back to interval 0 (enqueued on the context stream) */
static int filter_reset_async(plom_ctx *c, plom_filter *f)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(256), 0, c->stream, f->df, 1);
    HIPCHK(hipGetLastError());
    return PLOM_OK;
}

/* bytes of the host image of a filter (device layout from off_img to the end) */
static size_t image_bytes(const layout *L) { return L->total - L->off_img; }

/* Create a filter.  The descriptor, zeroed counters and converted inputs are written into
 * a host image that mirrors the device tail of the allocation, then uploaded in one copy.
 * With `pinned` (at least image_bytes, page-locked, kept by the caller until the stream
 * has passed the upload) the allocation and the copy are stream-ordered and nothing
 * waits; without it the call returns when the filter is on the device. */
/* Validation and host-side preparation of a filter's time-varying covariates (their table
 * and, for the ode implementation, their splines) and drift map; shared by the filter and
 * the server, which checks a request when it is submitted. */
struct input_prep {
    std::vector<double> covtab;             /* plom_cov_table, or empty */
    int cov_kmax = 0, cov_on = 0, drift_on = 0;
    std::vector<int32_t> dng;               /* groups per drifting parameter */
    std::vector<int32_t> spl_n, spl_off;    /* plom_spline_pack (ode with covariates) */
    std::vector<double> spl_val;
};

static int prepare_inputs(const plom_filter_desc *d, const plom_model_info &mi, const plom_filter_inputs *in,
                          input_prep *p)
{
    p->cov_on = in->cov != nullptr && mi.n_fix > 0;
    p->drift_on = mi.n_drift > 0 && !(d->noises_off & PLOM_NO_DRIFT_BIT) && !(d->flags & PLOM_FLAG_ODE);
    const plom_drift_map *dmap = in->drift;
    const size_t ND = (size_t)mi.n_drift * (size_t)mi.n_cac;
    p->dng.assign((size_t)mi.n_drift + 1, 0);
    if (p->drift_on) {
        if (!dmap || !dmap->group || !dmap->tr || !dmap->scale || !dmap->duration || !dmap->min || !dmap->max ||
            !dmap->multiplier || !dmap->vol_slot)
            return set_err(PLOM_ERR_ARG, "this model has drifting parameters: it needs a complete plom_drift_map "
                                         "(or PLOM_NO_DRIFT_BIT in noises_off)");
        for (size_t q = 0; q < ND; q++) {
            if (dmap->group[q] < 0 || dmap->group[q] >= mi.n_cac) return set_err(PLOM_ERR_ARG, "drift group[%zu] out of range", q);
            if (dmap->tr[q] < 0 || dmap->tr[q] > 3 || dmap->scale[q] < 0 || dmap->scale[q] > 2)
                return set_err(PLOM_ERR_ARG, "drift %zu: bad transform or scale", q);
            if (dmap->vol_slot[q] < -1 || dmap->vol_slot[q] >= mi.n_pp)
                return set_err(PLOM_ERR_ARG, "drift vol_slot[%zu]=%d is not a process parameter slot of pv", q, dmap->vol_slot[q]);
            const size_t i = q / (size_t)mi.n_cac;
            if (dmap->group[q] + 1 > p->dng[i]) p->dng[i] = dmap->group[q] + 1;
        }
    }
    /* time-varying covariates: their values at every substep and observation time; the ode
     * implementation also evaluates their splines itself, at its own times */
    if (p->cov_on) {
        p->cov_kmax = plom_cov_kmax(d->N, in->times, d->dt);
        p->covtab.resize(plom_cov_rows(d->N, p->cov_kmax) * (size_t)mi.n_fix);
        char msg[256];
        if (plom_cov_table(in->cov, mi.n_fix, d->N, in->times, d->dt, p->cov_kmax, p->covtab.data(), msg, sizeof msg))
            return set_err(PLOM_ERR_ARG, "%s", msg);
        if (d->flags & PLOM_FLAG_ODE) {
            p->spl_n.resize((size_t)mi.n_fix);
            p->spl_off.resize((size_t)mi.n_fix);
            p->spl_val.resize(plom_spline_vals(in->cov, mi.n_fix));
            if (plom_spline_pack(in->cov, mi.n_fix, p->spl_n.data(), p->spl_off.data(), p->spl_val.data(), msg, sizeof msg))
                return set_err(PLOM_ERR_ARG, "%s", msg);
        }
    }
    return PLOM_OK;
}

/* fixed: the device memory to lay the filter out in (the fixed scratch, fixed_cap bytes)
 * instead of an allocation of its own */
static int filter_create_impl(plom_ctx *c, const plom_filter_desc *d, const plom_filter_inputs *in, char *pinned,
                              plom_filter **out, hipStream_t st = nullptr, char *fixed = nullptr, size_t fixed_cap = 0)
{
    if (!st && c) st = c->stream;
    if (!c || !out || !in) return set_err(PLOM_ERR_ARG, "null argument");
    plom_model_info mi;
    int rc = check_desc(d, &mi);
    if (rc != PLOM_OK) return rc;
    if (!in->times || !in->pv || !in->x0 || (d->N > 0 && !in->data)) return set_err(PLOM_ERR_ARG, "null input array");
    const plom_mif_map *mm = in->mif;
    if (d->flags & PLOM_FLAG_MIF) {
        if (!mm || !mm->pv_src || !mm->tr || !mm->scale || !mm->duration || !mm->min || !mm->max || !mm->multiplier || !mm->sd)
            return set_err(PLOM_ERR_ARG, "PLOM_FLAG_MIF needs a complete plom_mif_map");
        for (int s2 = 0; s2 < mi.n_pp; s2++)
            if (mm->pv_src[s2] < -1 || mm->pv_src[s2] >= d->n_theta)
                return set_err(PLOM_ERR_ARG, "mif pv_src[%d]=%d outside [-1, n_theta)", s2, mm->pv_src[s2]);
        for (int o = 0; o < d->n_theta; o++)
            if (mm->tr[o] < 0 || mm->tr[o] > 3 || mm->scale[o] < 0 || mm->scale[o] > 2 || !(mm->sd[o] >= 0.0))
                return set_err(PLOM_ERR_ARG, "mif component %d: bad transform, scale or sd", o);
        if (mm->chol && d->n_theta > PLOM_MIF_MVN_MAX)
            return set_err(PLOM_ERR_ARG, "a correlated random walk (mif chol) takes at most %d components", PLOM_MIF_MVN_MAX);
        if ((mm->prior != nullptr) != (mm->prior_pow != nullptr) || (mm->prior != nullptr) != (mm->prior_until != nullptr))
            return set_err(PLOM_ERR_ARG, "mif prior, prior_pow and prior_until go together");
        for (int o = 0; mm->prior && o < d->n_theta; o++) {
            if (mm->prior[o] < PLOM_PRIOR_NONE || mm->prior[o] > PLOM_PRIOR_PSEUDO_UNIFORM)
                return set_err(PLOM_ERR_ARG, "mif component %d: unknown prior %d", o, mm->prior[o]);
            if (mm->prior[o] != PLOM_PRIOR_NONE && (!(mm->max[o] > mm->min[o]) || !(mm->prior_pow[o] >= 0.0) ||
                                                    mm->prior_pow[o] == PLOM_INF))
                return set_err(PLOM_ERR_ARG, "mif component %d: a prior needs min < max and a finite prior_pow >= 0", o);
        }
    }
    for (int n = 0; n < d->N; n++)
        if (!(in->times[n + 1] > in->times[n])) return set_err(PLOM_ERR_ARG, "times must increase (n=%d)", n);
    input_prep prep;
    rc = prepare_inputs(d, mi, in, &prep);
    if (rc != PLOM_OK) return rc;
    const std::vector<double> &covtab = prep.covtab;
    const int cov_kmax = prep.cov_kmax, cov_on = prep.cov_on, drift_on = prep.drift_on;
    const plom_drift_map *dmap = in->drift;
    const size_t ND = (size_t)mi.n_drift * (size_t)mi.n_cac;
    const std::vector<int32_t> &dng = prep.dng, &spl_n = prep.spl_n, &spl_off = prep.spl_off;
    const std::vector<double> &spl_val = prep.spl_val;
    HIPCHK(hipSetDevice(c->device));
    layout L;
    const size_t chol_vals = (d->flags & PLOM_FLAG_MIF) && mm && mm->chol ? (size_t)d->n_theta * d->n_theta : 0;
    const int prior_on = (d->flags & PLOM_FLAG_MIF) && mm && mm->prior;
    make_layout(d, &mi, &L, covtab.size(), chol_vals, prior_on, spl_val.size());
    const int async = pinned != nullptr;
    std::vector<char> pageable;
    char *img = pinned;
    if (!img) {
        pageable.assign(image_bytes(&L), 0);
        img = pageable.data();
    } else {
        memset(img, 0, image_bytes(&L));
    }
    plom_filter *f = new (std::nothrow) plom_filter();
    if (!f) return set_err(PLOM_ERR_NOMEM, "host allocation failed");
    f->desc = *d;
    f->info = mi;
    f->variant = (d->flags & PLOM_FLAG_ODE) ? 6
                 : (d->flags & PLOM_FLAG_PSR) ? 4 + ((d->noises_off & PLOM_NO_ENV_STO_BIT) ? 1 : 0)
                 : (int)(d->noises_off & 3u);
    f->dev_bytes = L.total;
    f->async_alloc = async;
    hipError_t e = hipSuccess;
    if (fixed) {
        if (L.total > fixed_cap) {
            delete f;
            return set_err(PLOM_ERR_ARG, "the filter needs %zu bytes of device memory, more than the %zu of the fixed scratch",
                           L.total, fixed_cap);
        }
        f->dev = fixed;
        f->fixed_dev = 1;
    } else {
        e = async ? hipMallocAsync((void **)&f->dev, L.total, st) : hipMalloc((void **)&f->dev, L.total);
        if (e != hipSuccess) {
            delete f;
            return set_err(PLOM_ERR_NOMEM, "hipMalloc(%zu): %s", L.total, hipGetErrorString(e));
        }
    }
    char *b = f->dev;
    plom_dfilter *h = &f->h;
    const int Jp = (d->J + TILE - 1) / TILE * TILE;
    h->J = d->J; h->Jp = Jp; h->T = Jp / TILE; h->N = d->N; h->S = mi.size_proj; h->n_ts = mi.n_ts;
    h->flags = (int32_t)d->flags; h->step = 0;
    h->seed = d->seed; h->stream = d->stream; h->prec = (uint32_t)d->precision;
    h->dt = d->dt;
    h->sqdt = plom_sqrt(d->dt);
    h->like_min = d->like_min;
    h->eps_abs = (d->eps_abs > 0.0 || d->eps_rel > 0.0) ? d->eps_abs : PLOM_ODE_EPS_ABS;
    h->eps_rel = (d->eps_abs > 0.0 || d->eps_rel > 0.0) ? d->eps_rel : PLOM_ODE_EPS_REL;
    h->all_fail_llike = plom_log(d->like_min) * (double)mi.n_ts;
    h->X[0] = b + L.off_X0; h->X[1] = b + L.off_X1;
    h->x0 = b + L.off_x0;
    h->pv = (const double *)(b + L.off_pv); h->pvf = (const float *)(b + L.off_pvf);
    h->times = (const double *)(b + L.off_times); h->data = (const double *)(b + L.off_data);
    h->nonan = (const int32_t *)(b + L.off_nonan);
    h->logw = (double *)(b + L.off_logw); h->tmax = (double *)(b + L.off_tmax);
    h->tsum = (double *)(b + L.off_tsum); h->tsum2 = (double *)(b + L.off_tsum2);
    h->cum = (uint64_t *)(b + L.off_cum); h->ttot = (uint64_t *)(b + L.off_ttot); h->tincl = (uint64_t *)(b + L.off_tincl);
    h->cnt = (uint32_t *)(b + L.off_cnt); h->scal = (double *)(b + L.off_scal);
    h->n_all_fail = (int32_t *)(b + L.off_naf);
    h->llike_n = (double *)(b + L.off_llike_n); h->ess_n = (double *)(b + L.off_ess_n);
    h->rflag = (int32_t *)(b + L.off_rflag);
    h->ancestry = (d->flags & (PLOM_FLAG_STORE_ANCESTRY | PLOM_FLAG_LINEAGE | PLOM_FLAG_ALL_STATES)) ? (int32_t *)(b + L.off_anc) : nullptr;
    h->allX = (d->flags & PLOM_FLAG_ALL_STATES) ? (void *)(b + L.off_allX) : nullptr;
    h->lin_slots = (d->flags & PLOM_FLAG_LINEAGE) ? (int32_t *)(b + L.off_lslots) : nullptr;
    h->lin_state = (d->flags & PLOM_FLAG_LINEAGE) ? (double *)(b + L.off_lstate) : nullptr;
    h->lin_obs = (d->flags & PLOM_FLAG_LINEAGE) ? (double *)(b + L.off_lobs) : nullptr;
    {
        int K, nseg;
        lineage_segments(d->N, d->flags, &K, &nseg);
        h->lin_K = K;
        h->lin_nseg = nseg;
        h->lin_snap = ((d->flags & PLOM_FLAG_LINEAGE) && nseg > 1) ? (void *)(b + L.off_lsnap) : nullptr;
    }
    const int hat_on = (d->flags & PLOM_FLAG_HAT) != 0;
    h->n_hat = hat_on ? mi.n_hat : 0;
    h->hkey = hat_on ? (uint64_t *)(b + L.off_hkey) : nullptr;
    h->hpref = hat_on ? (uint64_t *)(b + L.off_hpref) : nullptr;
    h->hpart = hat_on ? (double *)(b + L.off_hpart) : nullptr;
    h->hat = hat_on ? (double *)(b + L.off_hat) : nullptr;
    h->pres = hat_on ? (double *)(b + L.off_pres) : nullptr;
    h->prw = hat_on ? (double *)(b + L.off_prw) : nullptr;
    h->prp = hat_on ? (double *)(b + L.off_prp) : nullptr;
    const int mif_on = (d->flags & PLOM_FLAG_MIF) != 0;
    h->n_theta = mif_on ? d->n_theta : 0;
    h->n_pp = mif_on ? mi.n_pp : 0;
    h->pm_src = mif_on ? (const int32_t *)(b + L.off_pm_src) : nullptr;
    h->pm_tr = mif_on ? (const int32_t *)(b + L.off_pm_tr) : nullptr;
    h->pm_sc = mif_on ? (const int32_t *)(b + L.off_pm_sc) : nullptr;
    h->pm_dur = mif_on ? (const int32_t *)(b + L.off_pm_dur) : nullptr;
    h->pm_min = mif_on ? (const double *)(b + L.off_pm_min) : nullptr;
    h->pm_max = mif_on ? (const double *)(b + L.off_pm_max) : nullptr;
    h->pm_mult = mif_on ? (const double *)(b + L.off_pm_mult) : nullptr;
    h->pm_sd = mif_on ? (const double *)(b + L.off_pm_sd) : nullptr;
    h->mut_fac = mif_on ? (const double *)(b + L.off_mutfac) : nullptr;
    h->pm_chol = (mif_on && mm->chol) ? (const double *)(b + L.off_pm_chol) : nullptr;
    h->pm_prior = prior_on ? (const int32_t *)(b + L.off_pm_prior) : nullptr;
    h->pm_ppow = prior_on ? (const double *)(b + L.off_pm_ppow) : nullptr;
    h->pm_puntil = prior_on ? (const int32_t *)(b + L.off_pm_puntil) : nullptr;
    h->th[0] = mif_on ? (double *)(b + L.off_th0) : nullptr;
    h->th[1] = mif_on ? (double *)(b + L.off_th1) : nullptr;
    h->mw = mif_on ? (double *)(b + L.off_mw) : nullptr;
    h->mpw = mif_on ? (double *)(b + L.off_mpw) : nullptr;
    h->mps = mif_on ? (double *)(b + L.off_mps) : nullptr;
    h->mpm = mif_on ? (double *)(b + L.off_mpm) : nullptr;
    h->mpq = mif_on ? (double *)(b + L.off_mpq) : nullptr;
    h->mmean = mif_on ? (double *)(b + L.off_mmean) : nullptr;
    h->mvar = mif_on ? (double *)(b + L.off_mvar) : nullptr;
    h->drift_on = drift_on;
    if (drift_on) {
        h->dv.group = (const int32_t *)(b + L.off_dm_group);
        h->dv.ng = (const int32_t *)(b + L.off_dm_ng);
        h->dv.tr = (const int32_t *)(b + L.off_dm_tr);
        h->dv.sc = (const int32_t *)(b + L.off_dm_sc);
        h->dv.dur = (const int32_t *)(b + L.off_dm_dur);
        h->dv.min = (const double *)(b + L.off_dm_min);
        h->dv.max = (const double *)(b + L.off_dm_max);
        h->dv.mult = (const double *)(b + L.off_dm_mult);
        h->dv.vol_slot = (const int32_t *)(b + L.off_dm_vol);
    }
    if (!spl_val.empty()) {
        h->spl.n = (const int32_t *)(b + L.off_spl_n);
        h->spl.off = (const int32_t *)(b + L.off_spl_off);
        h->spl.val = (const double *)(b + L.off_spl_val);
        h->spl_on = 1;
    }
    if (cov_on) {
        h->cov = (const double *)(b + L.off_cov);
        h->cov_cs = mi.n_fix;
        h->cov_kmax = cov_kmax;
    } else {
        h->cov = h->pv + mi.fix_off;   /* one row, repeated: the constants in pv */
        h->cov_cs = 0;
        h->cov_kmax = 0;
    }
    f->df = (plom_dfilter *)(b + L.off_df);
    f->off_pv = L.off_pv; f->off_pvf = L.off_pvf; f->off_x0 = L.off_x0;
    if (d->flags & PLOM_FLAG_ALL_STATES) {
        f->pv_h.assign(in->pv, in->pv + mi.pv_size);
        f->times_h.assign(in->times, in->times + d->N + 1);
        if (cov_on) f->cov_h = covtab;
        f->cov_kmax_h = cov_kmax;
    }
    f->off_mutfac = L.off_mutfac;

    /* host image: descriptor, zeroed counters and outputs, inputs */
    const int S = mi.size_proj;
    auto at = [&](size_t off) { return img + (off - L.off_img); };  /* device offset -> image byte */
    memcpy(at(L.off_df), h, sizeof(plom_dfilter));
    int32_t *nonan = (int32_t *)at(L.off_nonan);
    f->nonan.assign((size_t)d->N, 0);
    for (int n = 0; n < d->N; n++) {
        int k = 0;
        for (int ts = 0; ts < mi.n_ts; ts++) k += !(in->data[(size_t)n * mi.n_ts + ts] != in->data[(size_t)n * mi.n_ts + ts]);
        nonan[n] = k;
        f->nonan[(size_t)n] = k;
    }
    if (mif_on) {
        memcpy(at(L.off_pm_src), mm->pv_src, (size_t)mi.n_pp * 4);
        memcpy(at(L.off_pm_tr), mm->tr, (size_t)d->n_theta * 4);
        memcpy(at(L.off_pm_sc), mm->scale, (size_t)d->n_theta * 4);
        memcpy(at(L.off_pm_dur), mm->duration, (size_t)d->n_theta * 4);
        memcpy(at(L.off_pm_min), mm->min, (size_t)d->n_theta * 8);
        memcpy(at(L.off_pm_max), mm->max, (size_t)d->n_theta * 8);
        memcpy(at(L.off_pm_mult), mm->multiplier, (size_t)d->n_theta * 8);
        memcpy(at(L.off_pm_sd), mm->sd, (size_t)d->n_theta * 8);
        if (mm->chol) memcpy(at(L.off_pm_chol), mm->chol, (size_t)d->n_theta * d->n_theta * 8);
        if (prior_on) {
            memcpy(at(L.off_pm_prior), mm->prior, (size_t)d->n_theta * 4);
            memcpy(at(L.off_pm_ppow), mm->prior_pow, (size_t)d->n_theta * 8);
            memcpy(at(L.off_pm_puntil), mm->prior_until, (size_t)d->n_theta * 4);
        }
    }
    memcpy(at(L.off_times), in->times, (size_t)(d->N + 1) * 8);
    if (d->N > 0) memcpy(at(L.off_data), in->data, (size_t)d->N * mi.n_ts * 8);
    memcpy(at(L.off_pv), in->pv, (size_t)mi.pv_size * 8);
    if (cov_on) memcpy(at(L.off_cov), covtab.data(), covtab.size() * 8);
    if (!spl_val.empty()) {
        memcpy(at(L.off_spl_n), spl_n.data(), spl_n.size() * 4);
        memcpy(at(L.off_spl_off), spl_off.data(), spl_off.size() * 4);
        memcpy(at(L.off_spl_val), spl_val.data(), spl_val.size() * 8);
    }
    if (drift_on) {
        memcpy(at(L.off_dm_group), dmap->group, ND * 4);
        memcpy(at(L.off_dm_ng), dng.data(), (size_t)mi.n_drift * 4);
        memcpy(at(L.off_dm_tr), dmap->tr, ND * 4);
        memcpy(at(L.off_dm_sc), dmap->scale, ND * 4);
        memcpy(at(L.off_dm_dur), dmap->duration, ND * 4);
        memcpy(at(L.off_dm_min), dmap->min, ND * 8);
        memcpy(at(L.off_dm_max), dmap->max, ND * 8);
        memcpy(at(L.off_dm_mult), dmap->multiplier, ND * 8);
        memcpy(at(L.off_dm_vol), dmap->vol_slot, ND * 4);
    }
    float *pvf = (float *)at(L.off_pvf);
    for (int i = 0; i < mi.pv_size; i++) pvf[i] = (float)in->pv[i];
    if (d->precision == PLOM_FP32) {
        float *x0 = (float *)at(L.off_x0);
        for (int k = 0; k < S; k++) x0[k] = (float)in->x0[k];
    } else {
        memcpy(at(L.off_x0), in->x0, (size_t)S * 8);
    }
    e = hipMemcpyAsync(b + L.off_img, img, image_bytes(&L), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !async) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        if (f->fixed_dev) (void)hipStreamSynchronize(st);
        else if (async) (void)hipFreeAsync(f->dev, st);
        else (void)hipFree(f->dev);
        delete f;
        return set_err(PLOM_ERR_HIP, "filter upload: %s", hipGetErrorString(e));
    }
    *out = f;
    return PLOM_OK;
}

extern "C" int plom_filter_create(plom_ctx *c, const plom_filter_desc *d, const plom_filter_inputs *in, plom_filter **out)
{
    return filter_create_impl(c, d, in, nullptr, out);
}

extern "C" int plom_filter_reset(plom_ctx *c, plom_filter *f)
{
    if (!c || !f) return set_err(PLOM_ERR_ARG, "null argument");
    f->has_run = 0;
    return filter_reset_async(c, f);
}

extern "C" int plom_filter_update(plom_ctx *c, plom_filter *f, const double *pv, const double *x0, uint32_t seed,
                                  uint32_t stream)
{
    if (!c || !f) return set_err(PLOM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    const int S = f->info.size_proj, P = f->info.pv_size;
    /* pv, pvf and x0 are consecutive regions of the image */
    std::vector<char> img(f->off_x0 + (size_t)S * 8 - f->off_pv, 0);
    char *base = img.data();
    if (!pv || !x0) {
        /* keep what is on the device for the part not given */
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(base, f->dev + f->off_pv, img.size(), hipMemcpyDeviceToHost));
    }
    if (pv) {
        memcpy(base, pv, (size_t)P * 8);
        if (!f->pv_h.empty()) f->pv_h.assign(pv, pv + P);
        float *pvf = (float *)(base + (f->off_pvf - f->off_pv));
        for (int i = 0; i < P; i++) pvf[i] = (float)pv[i];
    }
    if (x0) {
        char *xd = base + (f->off_x0 - f->off_pv);
        if (f->desc.precision == PLOM_FP32) {
            for (int k = 0; k < S; k++) ((float *)xd)[k] = (float)x0[k];
        } else {
            memcpy(xd, x0, (size_t)S * 8);
        }
    }
    HIPCHK(hipMemcpyAsync(f->dev + f->off_pv, base, img.size(), hipMemcpyHostToDevice, c->stream));
    f->desc.seed = seed;
    f->desc.stream = stream;
    f->h.seed = seed;
    f->h.stream = stream;
    HIPCHK(hipMemcpyAsync(f->df, &f->h, sizeof(plom_dfilter), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    f->dirty = 1;   /* has_run stays: the next run starts with a reset */
    return PLOM_OK;
}

extern "C" int plom_filter_mif_start(plom_ctx *c, plom_filter *f, const double *theta, const double *x0,
                                     const double *mut_fac, uint32_t seed, uint32_t stream)
{
    if (!c || !f || !theta || !x0 || (f->desc.N > 0 && !mut_fac)) return set_err(PLOM_ERR_ARG, "null argument");
    if (!(f->desc.flags & PLOM_FLAG_MIF)) return set_err(PLOM_ERR_STATE, "filter was not created with PLOM_FLAG_MIF");
    HIPCHK(hipSetDevice(c->device));
    const size_t J = (size_t)f->desc.J, Jp = (size_t)f->h.Jp, S = (size_t)f->info.size_proj, NT = (size_t)f->desc.n_theta;
    const size_t N = (size_t)f->desc.N;
    for (size_t n = 0; n < N; n++)
        if (!(mut_fac[n] >= 0.0)) return set_err(PLOM_ERR_ARG, "mut_fac[%zu] must be >= 0", n);
    /* theta -> th[0] and x0 -> X[0], padded rows, in one staged copy each */
    std::vector<double> tb(NT * Jp, 0.0);
    for (size_t o = 0; o < NT; o++) memcpy(tb.data() + o * Jp, theta + o * J, J * 8);
    HIPCHK(hipMemcpyAsync(f->h.th[0], tb.data(), tb.size() * 8, hipMemcpyHostToDevice, c->stream));
    std::vector<double> xb;
    std::vector<float> xf;
    if (f->desc.precision == PLOM_FP32) {
        xf.assign(S * Jp, 0.0f);
        for (size_t k = 0; k < S; k++)
            for (size_t j = 0; j < J; j++) xf[k * Jp + j] = (float)x0[k * J + j];
        HIPCHK(hipMemcpyAsync(f->h.X[0], xf.data(), xf.size() * 4, hipMemcpyHostToDevice, c->stream));
    } else {
        xb.assign(S * Jp, 0.0);
        for (size_t k = 0; k < S; k++) memcpy(xb.data() + k * Jp, x0 + k * J, J * 8);
        HIPCHK(hipMemcpyAsync(f->h.X[0], xb.data(), xb.size() * 8, hipMemcpyHostToDevice, c->stream));
    }
    if (N) HIPCHK(hipMemcpyAsync(f->dev + f->off_mutfac, mut_fac, N * 8, hipMemcpyHostToDevice, c->stream));
    f->desc.seed = seed;
    f->desc.stream = stream;
    f->h.seed = seed;
    f->h.stream = stream;
    HIPCHK(hipMemcpyAsync(f->df, &f->h, sizeof(plom_dfilter), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    f->dirty = 1;
    f->mif_ready = 1;
    return PLOM_OK;
}

extern "C" int plom_filter_destroy(plom_ctx *c, plom_filter *f)
{
    if (!f) return PLOM_OK;
    if (c) HIPCHK(hipStreamSynchronize(c->stream));
    if (f->fixed_dev) { /* the fixed scratch stays */ }
    else if (f->async_alloc && c) HIPCHK(hipFreeAsync(f->dev, c->stream));
    else HIPCHK(hipFree(f->dev));
    delete f;
    return PLOM_OK;
}

/* A side stream for the theta moments of MIF filters (k_mif_tiles, k_mif_reduce): the
 * moments of interval n run beside interval n + 1's propagate instead of before it.  They
 * read the scan's copy of the weights (mw), the interval's flags and the theta table the
 * next propagate only reads, and write buffers nothing else touches; the next scan, which
 * rewrites mw, waits for them.  `pending`: moments enqueued and not yet waited for. */
struct mif_side {
    hipStream_t st;
    hipEvent_t fork, done;
    int pending;
};

static int ctx_mif_side(plom_ctx *c, mif_side *sd)
{
    if (!c->mside) {
        HIPCHK(hipStreamCreateWithFlags(&c->mside, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&c->mside_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->mside_done, hipEventDisableTiming));
    }
    sd->st = c->mside;
    sd->fork = c->mside_fork;
    sd->done = c->mside_done;
    sd->pending = 0;
    return PLOM_OK;
}

/* the stream s waits for the moments on the side stream */
static int mif_side_join(hipStream_t s, mif_side *sd)
{
    if (!sd || !sd->pending) return PLOM_OK;
    HIPCHK(hipStreamWaitEvent(s, sd->done, 0));
    sd->pending = 0;
    return PLOM_OK;
}

/* weights, fixed-point scan and tile prefix of interval n: K2 (one block per filter), K3
 * (one block per tile), K4 (one block per filter); with MIF filters (mif_theta, the largest
 * n_theta of them, > 0) the theta moments of the weighted cloud after the scan, on the
 * stream s or, with a side stream, beside the next interval */
static int enqueue_resample(hipStream_t s, plom_dfilter *df, const int2 *tiles, int ntiles, const int *fidx, int nfilters,
                            int mif_theta = 0, int n = 0, mif_side *side = nullptr)
{
    hipLaunchKernelGGL(k_weights, dim3(nfilters), dim3(TILE), 0, s, df, fidx);
    if (mif_theta > 0) {
        const dim3 gt(ntiles, (mif_theta + MIF_COMPS - 1) / MIF_COMPS), gr(nfilters, mif_theta);
        if (side) {
            if (mif_side_join(s, side) != PLOM_OK) return PLOM_ERR_HIP;   /* k_scan rewrites mw */
            hipLaunchKernelGGL(k_scan, dim3(ntiles), dim3(TILE), 0, s, df, tiles);
            hipLaunchKernelGGL(k_advance, dim3(nfilters), dim3(TILE), 0, s, df, fidx);
            HIPCHK(hipEventRecord(side->fork, s));
            HIPCHK(hipStreamWaitEvent(side->st, side->fork, 0));
            hipLaunchKernelGGL(k_mif_tiles, gt, dim3(TILE), 0, side->st, df, tiles, n);
            hipLaunchKernelGGL(k_mif_reduce, gr, dim3(TILE), 0, side->st, df, fidx, n);
            HIPCHK(hipEventRecord(side->done, side->st));
            side->pending = 1;
            return PLOM_OK;
        }
        hipLaunchKernelGGL(k_scan, dim3(ntiles), dim3(TILE), 0, s, df, tiles);
        hipLaunchKernelGGL(k_mif_tiles, gt, dim3(TILE), 0, s, df, tiles, n);
        hipLaunchKernelGGL(k_mif_reduce, gr, dim3(TILE), 0, s, df, fidx, n);
        hipLaunchKernelGGL(k_advance, dim3(nfilters), dim3(TILE), 0, s, df, fidx);
        return PLOM_OK;
    }
    hipLaunchKernelGGL(k_scan, dim3(ntiles), dim3(TILE), 0, s, df, tiles);
    hipLaunchKernelGGL(k_advance, dim3(nfilters), dim3(TILE), 0, s, df, fidx);
    return PLOM_OK;
}

static void enqueue_final(plom_ctx *c, plom_filter *f, const int2 *tiles, plom_dfilter *df, int nblocks)
{
    if (f->desc.precision == PLOM_FP32)
        hipLaunchKernelGGL((k_final<float>), dim3(nblocks), dim3(TILE), 0, c->stream, df, tiles, f->info.size_proj);
    else
        hipLaunchKernelGGL((k_final<double>), dim3(nblocks), dim3(TILE), 0, c->stream, df, tiles, f->info.size_proj);
}

static int needs_final(const plom_filter *f)
{
    return f->desc.N > 0 && (f->desc.flags & (PLOM_FLAG_STORE_ANCESTRY | PLOM_FLAG_FINAL_STATE | PLOM_FLAG_ALL_STATES));
}

extern "C" int plom_filter_run(plom_ctx *c, plom_filter *f)
{
    if (!c || !f) return set_err(PLOM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    if (f->has_run) {
        int rc = filter_reset_async(c, f);
        if (rc != PLOM_OK) return rc;
    }
    const model_entry *me = model_at(f->desc.model_id);
    const int T = f->h.T;
    const int mif = (f->desc.flags & PLOM_FLAG_MIF) != 0;
    if (mif) {
        if (!f->mif_ready) return set_err(PLOM_ERR_STATE, "PLOM_FLAG_MIF: call plom_filter_mif_start before every run");
        f->mif_ready = 0;
    }
    (void)hipGetLastError(); /* a not-ready query or an unused event earlier must not be reported as a launch failure */
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    int k = 0;
    c->kprop_launches = 0;
    c->kspans.clear();
    c->kspan_rounds = c->ktiming ? f->desc.N : 0;
    mif_side side;
    if (mif && ctx_mif_side(c, &side) != PLOM_OK) return PLOM_ERR_HIP;
    for (int n = 0; n < f->desc.N; n++) {
        const int a = kspan_open(c, &k, c->stream);
        me->launch_prop(f->desc.precision, f->variant, mif, T, c->tpb, c->stream, f->df, nullptr);
        c->kprop_launches++;
        const int b = kspan_close(c, &k, c->stream, a, 0, 0);
        if (enqueue_resample(c->stream, f->df, nullptr, T, nullptr, 1, mif ? f->desc.n_theta : 0, n, mif ? &side : nullptr) != PLOM_OK)
            return PLOM_ERR_HIP;
        if (kspan_close(c, &k, c->stream, b, 1, 0) == -2) return PLOM_ERR_HIP;
        if (f->desc.flags & PLOM_FLAG_HAT) launch_hat_all(me, f->desc.precision, c->stream, f->df, T, f->info.n_hat);
    }
    if (mif && mif_side_join(c->stream, &side) != PLOM_OK) return PLOM_ERR_HIP;
    if ((f->desc.flags & PLOM_FLAG_HAT) && f->desc.N > 0) me->launch_hat0(f->desc.precision, c->stream, f->df);
    if (needs_final(f)) enqueue_final(c, f, nullptr, f->df, T);
    if (f->desc.N > 0 && (f->desc.flags & PLOM_FLAG_LINEAGE))
        me->launch_lineage(f->desc.precision, f->variant, c->stream, f->df);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    f->has_run = 1;
    return PLOM_OK;
}

/* copy a finished filter's results to the host on stream `s` (the caller has made sure
 * that the filter's work is complete) */
static int filter_read_on(plom_ctx *c, plom_filter *f, plom_filter_outputs *o, hipStream_t s)
{
    const plom_dfilter *h = &f->h;
    const size_t N = (size_t)f->desc.N, J = (size_t)f->desc.J;
    if (o->ancestry && N && !h->ancestry) return set_err(PLOM_ERR_STATE, "ancestry requested but PLOM_FLAG_STORE_ANCESTRY not set");
    if (o->final_state && (!(f->desc.flags & PLOM_FLAG_FINAL_STATE) || N == 0))
        return set_err(PLOM_ERR_STATE, "final state requested but PLOM_FLAG_FINAL_STATE not set (or N == 0)");
    double scal[4];
    int32_t naf = 0;
    HIPCHK(hipMemcpyAsync(scal, h->scal, sizeof scal, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&naf, h->n_all_fail, sizeof naf, hipMemcpyDeviceToHost, s));
    if (o->loglik_n && N) HIPCHK(hipMemcpyAsync(o->loglik_n, h->llike_n, N * 8, hipMemcpyDeviceToHost, s));
    if (o->ess_n && N) HIPCHK(hipMemcpyAsync(o->ess_n, h->ess_n, N * 8, hipMemcpyDeviceToHost, s));
    if (o->ancestry && N) HIPCHK(hipMemcpyAsync(o->ancestry, h->ancestry, N * J * 4, hipMemcpyDeviceToHost, s));
    if (o->final_logw) HIPCHK(hipMemcpyAsync(o->final_logw, h->logw, J * 8, hipMemcpyDeviceToHost, s));
    if ((o->hat || o->pred_res) && N) {
        if (!h->hat) return set_err(PLOM_ERR_STATE, "hat requested but PLOM_FLAG_HAT not set");
        if (o->hat) HIPCHK(hipMemcpyAsync(o->hat, h->hat, (N + 1) * (size_t)h->n_hat * 3 * 8, hipMemcpyDeviceToHost, s));
        if (o->pred_res) HIPCHK(hipMemcpyAsync(o->pred_res, h->pres, N * (size_t)h->n_ts * 2 * 8, hipMemcpyDeviceToHost, s));
    }
    if ((o->lineage || o->lineage_slots || o->lineage_obs) && N) {
        if (!h->lin_state) return set_err(PLOM_ERR_STATE, "lineage requested but PLOM_FLAG_LINEAGE not set");
        if (o->lineage_obs)
            HIPCHK(hipMemcpyAsync(o->lineage_obs, h->lin_obs, (N + 1) * (size_t)h->n_ts * 2 * 8, hipMemcpyDeviceToHost, s));
        if (o->lineage)
            HIPCHK(hipMemcpyAsync(o->lineage, h->lin_state, (N + 1) * (size_t)f->info.size_proj * 8, hipMemcpyDeviceToHost, s));
        if (o->lineage_slots) HIPCHK(hipMemcpyAsync(o->lineage_slots, h->lin_slots, N * 4, hipMemcpyDeviceToHost, s));
    }
    if ((o->mif_mean || o->mif_var) && !h->mmean) return set_err(PLOM_ERR_STATE, "mif moments requested but PLOM_FLAG_MIF not set");
    const size_t NT = (size_t)h->n_theta;
    if (o->mif_mean && N) HIPCHK(hipMemcpyAsync(o->mif_mean + NT, h->mmean + NT, N * NT * 8, hipMemcpyDeviceToHost, s));
    if (o->mif_var && N) HIPCHK(hipMemcpyAsync(o->mif_var + NT, h->mvar + NT, N * NT * 8, hipMemcpyDeviceToHost, s));
    std::vector<float> tmp;
    if (o->final_state) {
        const int S = f->info.size_proj;
        const void *src = h->X[(N + 1) & 1];
        if (f->desc.precision == PLOM_FP32) {
            tmp.resize((size_t)S * J);
            for (int k = 0; k < S; k++)
                HIPCHK(hipMemcpyAsync(tmp.data() + (size_t)k * J, (const float *)src + (size_t)k * h->Jp, J * 4, hipMemcpyDeviceToHost, s));
        } else {
            for (int k = 0; k < S; k++)
                HIPCHK(hipMemcpyAsync(o->final_state + (size_t)k * J, (const double *)src + (size_t)k * h->Jp, J * 8, hipMemcpyDeviceToHost, s));
        }
    }
    /* the stored clouds: the propagated cloud of every interval and the ancestry */
    const int want_all = o->all_states || o->all_obs;
    std::vector<double> allX;
    std::vector<float> allXf;
    std::vector<int32_t> anc;
    const size_t S = (size_t)f->info.size_proj, Jp = (size_t)h->Jp;
    if (want_all) {
        if (!h->allX) return set_err(PLOM_ERR_STATE, "all states requested but PLOM_FLAG_ALL_STATES not set");
        if (f->desc.precision == PLOM_FP32) {
            allXf.resize((N + 1) * S * Jp);
            HIPCHK(hipMemcpyAsync(allXf.data(), h->allX, allXf.size() * 4, hipMemcpyDeviceToHost, s));
        } else {
            allX.resize((N + 1) * S * Jp);
            HIPCHK(hipMemcpyAsync(allX.data(), h->allX, allX.size() * 8, hipMemcpyDeviceToHost, s));
        }
        anc.resize(N * J + 1);
        if (N) HIPCHK(hipMemcpyAsync(anc.data(), h->ancestry, N * J * 4, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    if (!tmp.empty())
        for (size_t i = 0; i < tmp.size(); i++) o->final_state[i] = tmp[i];
    if (want_all) {
        /* row n = the cloud after interval n-1 as resampling leaves it: particle j holds its
         * ancestor's state (identity where the interval was not resampled) */
        const model_entry *me = model_at(f->desc.model_id);
        const int NTS = f->info.n_ts;
        std::vector<double> x(S), obs((size_t)NTS + 1), mean((size_t)NTS + 1), var((size_t)NTS + 1);
        for (size_t n = 0; n <= N; n++) {
            const double *cvd = f->cov_h.empty()
                ? f->pv_h.data() + f->info.fix_off
                : f->cov_h.data() + (N * (size_t)f->cov_kmax_h + n) * (size_t)f->info.n_fix;
            for (size_t j = 0; j < J; j++) {
                const size_t src = n ? (size_t)anc[(n - 1) * J + j] : j;
                for (size_t k = 0; k < S; k++) {
                    const size_t at = (n * S + k) * Jp + src;
                    x[k] = allX.empty() ? (double)allXf[at] : allX[at];
                }
                if (o->all_states) memcpy(o->all_states + (n * J + j) * S, x.data(), S * 8);
                if (o->all_obs) {
                    me->proj2obs_h(x.data(), f->pv_h.data(), cvd, obs.data());
                    me->moments_h(obs.data(), f->pv_h.data(), cvd, f->times_h[n], mean.data(), var.data());
                    plom_obs_draw(f->desc.seed, f->desc.stream, (uint32_t)n, (uint32_t)(j + 1), NTS, mean.data(), var.data(),
                                  g_bm_tab_host, o->all_obs + (n * J + j) * (size_t)NTS * 2);
                }
            }
        }
    }
    /* rows without data (and row 0) carry no moments */
    for (size_t r = 0; r <= N && (o->mif_mean || o->mif_var) && NT; r++)
        if (r == 0 || f->nonan[r - 1] == 0)
            for (size_t q = 0; q < NT; q++) {
                if (o->mif_mean) o->mif_mean[r * NT + q] = NAN;
                if (o->mif_var) o->mif_var[r * NT + q] = NAN;
            }
    o->loglik = scal[3];
    o->n_all_fail = naf;
    return PLOM_OK;
}

extern "C" int plom_filter_read(plom_ctx *c, plom_filter *f, plom_filter_outputs *o)
{
    if (!c || !f || !o) return set_err(PLOM_ERR_ARG, "null argument");
    if (!f->has_run) return set_err(PLOM_ERR_STATE, "filter has not run");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return filter_read_on(c, f, o, c->stream);
}

extern "C" int plom_smc_run(plom_ctx *c, const plom_filter_desc *d, const plom_filter_inputs *in, plom_filter_outputs *o)
{
    plom_filter *f = nullptr;
    int rc = plom_filter_create(c, d, in, &f);
    if (rc != PLOM_OK) return rc;
    rc = plom_filter_run(c, f);
    if (rc == PLOM_OK) rc = plom_filter_read(c, f, o);
    int rc2 = plom_filter_destroy(c, f);
    return rc != PLOM_OK ? rc : rc2;
}

/* plom_smc_run in the device's fixed scratch: no device allocation or free.  The filter's
 * whole device layout (work buffers, counters, the uploaded inputs) is written before it is
 * read: the upload carries the zeroed counters and the inputs, and every interval writes the
 * states, weights and tile sums before reading them, so the previous call's contents in the
 * scratch do not reach the results. */
extern "C" int plom_smc_run_noalloc(plom_ctx *c, const plom_filter_desc *d, const plom_filter_inputs *in,
                                    plom_filter_outputs *o)
{
    if (!c || !d || !in || !o) return set_err(PLOM_ERR_ARG, "null argument");
    if (!c->scratch) return set_err(PLOM_ERR_ARG, "device %d has no fixed scratch", c->device);
    HIPCHK(hipSetDevice(c->device));
    plom_filter *f = nullptr;
    int rc = filter_create_impl(c, d, in, nullptr, &f, c->stream, c->scratch, PLOM_NOALLOC_MAX_BYTES);
    if (rc != PLOM_OK) return rc;
    rc = plom_filter_run(c, f);
    if (rc == PLOM_OK) rc = plom_filter_read(c, f, o);
    const int rc2 = plom_filter_destroy(c, f);
    return rc != PLOM_OK ? rc : rc2;
}

/* ------------------------------------------------------------------ batches */
/* A batch plan splits its filters into `lanes` chains (round-robin over the N-descending
 * order, so the chains carry similar work).  Each chain advances its filters together, one
 * launch round per interval, on a stream of its own; the chains fork from and join back
 * into the context stream, also inside a captured graph.  Two or more chains let one
 * chain's blocks fill the SIMDs that another leaves idle at the end of a propagate launch
 * and run beside its small weight / scan / advance launches.  Filters never interact, so
 * the results do not depend on the number of chains. */
struct batch_chain {
    std::vector<plom_filter *> fs;     /* N-descending: row r of the chain's descriptor array */
    struct group { int model, prec, variant, mif; std::vector<int2> tiles; int2 *d; int tpb; };
    std::vector<group> groups;
    std::vector<int2> all_tiles;
    int2 *d_all = nullptr;
    plom_dfilter *ddf = nullptr;
    char *buf = nullptr;
    int maxN = 0;
    int any_mif = 0;                   /* largest n_theta of the chain's MIF filters (0: none) */
};

/* tiles per propagate block chosen for the groups of one (model, precision, variant, MIF) */
struct tpb_rule { int model, prec, variant, mif, tpb; };

struct plom_batch {
    int nf;
    std::vector<plom_filter *> fs;     /* all filters, N-descending */
    int lanes;
    std::vector<batch_chain> chains;
    hipGraphExec_t graph;
    std::vector<tpb_rule> tpb_rules;   /* set by plom_batch_tune; other groups take the context's */
};

/* active tiles at interval s: a chain's filters are sorted by N, so the active ones are a prefix */
static int prefix_len(const batch_chain *ch, const std::vector<int2> &tl, int s)
{
    int k = 0;
    while (k < (int)tl.size() && ch->fs[(size_t)tl[(size_t)k].x]->desc.N > s) k++;
    return k;
}

static int chain_build(plom_ctx *c, batch_chain *ch)
{
    const int nf = (int)ch->fs.size();
    ch->maxN = ch->fs[0]->desc.N;
    for (int r = 0; r < nf; r++) {
        plom_filter *f = ch->fs[(size_t)r];
        batch_chain::group *g = nullptr;
        const int mif = (f->desc.flags & PLOM_FLAG_MIF) != 0;
        if (mif && f->desc.n_theta > ch->any_mif) ch->any_mif = f->desc.n_theta;
        for (auto &gg : ch->groups)
            if (gg.model == f->desc.model_id && gg.prec == f->desc.precision && gg.variant == f->variant && gg.mif == mif) g = &gg;
        if (!g) {
            ch->groups.push_back(batch_chain::group{f->desc.model_id, f->desc.precision, f->variant, mif, {}, nullptr, 1});
            g = &ch->groups.back();
        }
        for (int t = 0; t < f->h.T; t++) {
            g->tiles.push_back(make_int2(r, t));
            ch->all_tiles.push_back(make_int2(r, t));
        }
    }
    size_t bytes = align256((size_t)nf * sizeof(plom_dfilter)) + align256(ch->all_tiles.size() * sizeof(int2));
    for (auto &g : ch->groups) bytes += align256(g.tiles.size() * sizeof(int2));
    hipError_t e = hipMalloc((void **)&ch->buf, bytes);
    if (e != hipSuccess) {
        ch->buf = nullptr;
        return set_err(PLOM_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    std::vector<plom_dfilter> hd((size_t)nf);
    for (int r = 0; r < nf; r++) hd[(size_t)r] = ch->fs[(size_t)r]->h;
    ch->ddf = (plom_dfilter *)ch->buf;
    size_t o = align256((size_t)nf * sizeof(plom_dfilter));
    ch->d_all = (int2 *)(ch->buf + o);
    o += align256(ch->all_tiles.size() * sizeof(int2));
    for (auto &g : ch->groups) { g.d = (int2 *)(ch->buf + o); o += align256(g.tiles.size() * sizeof(int2)); }
    HIPCHK(hipMemcpyAsync(ch->ddf, hd.data(), (size_t)nf * sizeof(plom_dfilter), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(ch->d_all, ch->all_tiles.data(), ch->all_tiles.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    for (auto &g : ch->groups)
        HIPCHK(hipMemcpyAsync(g.d, g.tiles.data(), g.tiles.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PLOM_OK;
}

static void chain_reset(batch_chain *ch, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset, dim3((unsigned)ch->fs.size()), dim3(256), 0, st, ch->ddf, (int)ch->fs.size());
}

static int chain_propagate(batch_chain *ch, int s, hipStream_t st)
{
    int launches = 0;
    for (auto &g : ch->groups) {
        const int nb = prefix_len(ch, g.tiles, s);
        if (nb > 0) { model_at(g.model)->launch_prop(g.prec, g.variant, g.mif, nb, g.tpb, st, ch->ddf, g.d); launches++; }
    }
    return launches;
}

static int chain_resample(batch_chain *ch, int s, hipStream_t st, mif_side *side = nullptr)
{
    const int nb = prefix_len(ch, ch->all_tiles, s);
    int nfa = 0;
    while (nfa < (int)ch->fs.size() && ch->fs[(size_t)nfa]->desc.N > s) nfa++;
    if (enqueue_resample(st, ch->ddf, (const int2 *)ch->d_all, nb, nullptr, nfa, ch->any_mif, s, ch->any_mif ? side : nullptr) != PLOM_OK)
        return PLOM_ERR_HIP;
    for (int r = 0; r < nfa; r++) {
        plom_filter *f = ch->fs[(size_t)r];
        if (f->desc.flags & PLOM_FLAG_HAT)
            launch_hat_all(model_at(f->desc.model_id), f->desc.precision, st, ch->ddf + r, f->h.T, f->info.n_hat);
    }
    return PLOM_OK;
}

static void chain_final(plom_ctx *c, batch_chain *ch, hipStream_t st)
{
    size_t first = 0;
    for (int r = 0; r < (int)ch->fs.size(); r++) {
        plom_filter *f = ch->fs[(size_t)r];
        while (ch->all_tiles[first].x != r) first++;
        if (f->desc.N > 0 && (f->desc.flags & PLOM_FLAG_LINEAGE))
            model_at(f->desc.model_id)->launch_lineage(f->desc.precision, f->variant, st, ch->ddf + r);
        if (f->desc.N > 0 && (f->desc.flags & PLOM_FLAG_HAT))
            model_at(f->desc.model_id)->launch_hat0(f->desc.precision, st, ch->ddf + r);
        if (!needs_final(f)) continue;
        if (f->desc.precision == PLOM_FP32)
            hipLaunchKernelGGL((k_final<float>), dim3(f->h.T), dim3(TILE), 0, st, ch->ddf, ch->d_all + first, f->info.size_proj);
        else
            hipLaunchKernelGGL((k_final<double>), dim3(f->h.T), dim3(TILE), 0, st, ch->ddf, ch->d_all + first, f->info.size_proj);
    }
    (void)c;
}

/* streams and events of the extra lanes: lane 0 is the context stream, lane l >= 1 is
 * c->lane[l - 1] with join event c->join_ev[l - 1] */
static int ctx_lanes(plom_ctx *c, int lanes)
{
    if (!c->fork_ev) HIPCHK(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
    while ((int)c->lane.size() < lanes - 1) {
        hipStream_t st;
        hipEvent_t ev;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        c->lane.push_back(st);
        c->join_ev.push_back(ev);
    }
    return PLOM_OK;
}

/* Destroy the extra lanes' streams and the MIF side stream (after their work).  The
 * device has a few hardware queues per process (4 by default) and streams beyond that
 * share them; a stream that shares the context stream's queue runs in its order, not
 * beside it.  They are released when a batch has been tuned and before a server creates
 * its heavy lane, so that the heavy lane gets a queue of its own; a later launch that
 * needs them creates them again. */
static int ctx_release_lanes(plom_ctx *c)
{
    for (size_t l = 0; l < c->lane.size(); l++) {
        HIPCHK(hipStreamSynchronize(c->lane[l]));
        HIPCHK(hipStreamDestroy(c->lane[l]));
        HIPCHK(hipEventDestroy(c->join_ev[l]));
    }
    c->lane.clear();
    c->join_ev.clear();
    if (c->mside) {
        HIPCHK(hipStreamSynchronize(c->mside));
        HIPCHK(hipStreamDestroy(c->mside));
        HIPCHK(hipEventDestroy(c->mside_fork));
        HIPCHK(hipEventDestroy(c->mside_done));
        c->mside = nullptr;
    }
    return PLOM_OK;
}

/* the launch sequence of a batch (reset, every interval, final resampling); with
 * rounds >= 0 (the autotuner) only `rounds` launch rounds whose grids are those of
 * intervals spread evenly over the pass (a round costs what its active tiles cost, the
 * interval the device is at does not matter) and no final resampling.  With one lane every
 * chain runs on the context stream, interleaved round by round; with several, each lane's
 * chain runs on that lane's stream.  With timing, the propagate launches and the weight /
 * scan launches of every round are bracketed by events on the stream they run on. */
static int batch_enqueue(plom_ctx *c, plom_batch *b, int timing, int rounds)
{
    (void)hipGetLastError(); /* a not-ready query or an unused event earlier must not be reported as a launch failure */
    const int maxN = b->fs[0]->desc.N;
    const int R = rounds >= 0 && rounds < maxN ? rounds : maxN;
    const int stride = R < maxN ? maxN / R : 1;
    c->kprop_launches = 0;
    c->kspans.clear();
    c->kspan_rounds = timing ? R : 0;
    const int ktiming = c->ktiming;
    c->ktiming = timing;
    int k = 0;
    if (b->lanes == 1) {
        /* one lane: MIF moments on the side stream, beside the next round's propagate */
        bool any_mif = false;
        for (auto &ch : b->chains) any_mif = any_mif || ch.any_mif > 0;
        mif_side side;
        if (any_mif && ctx_mif_side(c, &side) != PLOM_OK) { c->ktiming = ktiming; return PLOM_ERR_HIP; }
        for (auto &ch : b->chains) chain_reset(&ch, c->stream);
        for (int i = 0; i < R; i++) {
            const int s = i * stride;
            const int a = kspan_open(c, &k, c->stream);
            for (auto &ch : b->chains)
                if (s < ch.maxN) c->kprop_launches += chain_propagate(&ch, s, c->stream);
            const int m = kspan_close(c, &k, c->stream, a, 0, 0);
            for (auto &ch : b->chains)
                if (s < ch.maxN && chain_resample(&ch, s, c->stream, any_mif ? &side : nullptr) != PLOM_OK) {
                    c->ktiming = ktiming;
                    return PLOM_ERR_HIP;
                }
            if (kspan_close(c, &k, c->stream, m, 1, 0) == -2) { c->ktiming = ktiming; return PLOM_ERR_HIP; }
        }
        if (any_mif && mif_side_join(c->stream, &side) != PLOM_OK) { c->ktiming = ktiming; return PLOM_ERR_HIP; }
        if (R == maxN)
            for (auto &ch : b->chains) chain_final(c, &ch, c->stream);
    } else {
        int rc = ctx_lanes(c, b->lanes);
        if (rc != PLOM_OK) return rc;
        HIPCHK(hipEventRecord(c->fork_ev, c->stream));
        for (int l = 1; l < b->lanes; l++) HIPCHK(hipStreamWaitEvent(c->lane[(size_t)l - 1], c->fork_ev, 0));
        for (int l = 0; l < b->lanes; l++) {
            batch_chain *ch = &b->chains[(size_t)l];
            hipStream_t st = l == 0 ? c->stream : c->lane[(size_t)l - 1];
            chain_reset(ch, st);
            /* with timing on, every chain brackets its own launches with events on its own
             * stream: the durations are those of the concurrent launches */
            for (int i = 0; i < R && i * stride < ch->maxN; i++) {
                const int a = kspan_open(c, &k, st);
                c->kprop_launches += chain_propagate(ch, i * stride, st);
                const int m = kspan_close(c, &k, st, a, 0, l);
                if (chain_resample(ch, i * stride, st) != PLOM_OK) { c->ktiming = ktiming; return PLOM_ERR_HIP; }
                if (kspan_close(c, &k, st, m, 1, l) == -2) { c->ktiming = ktiming; return PLOM_ERR_HIP; }
            }
            if (R == maxN) chain_final(c, ch, st);
        }
        for (int l = 1; l < b->lanes; l++) {
            HIPCHK(hipEventRecord(c->join_ev[(size_t)l - 1], c->lane[(size_t)l - 1]));
            HIPCHK(hipStreamWaitEvent(c->stream, c->join_ev[(size_t)l - 1], 0));
        }
    }
    c->ktiming = ktiming;
    HIPCHK(hipGetLastError());
    return PLOM_OK;
}

static int batch_launch(plom_ctx *c, plom_batch *b, int use_graph)
{
    for (auto f : b->fs)
        if ((f->desc.flags & PLOM_FLAG_MIF) && !f->mif_ready)
            return set_err(PLOM_ERR_STATE, "PLOM_FLAG_MIF: call plom_filter_mif_start before every launch");
    for (auto f : b->fs) f->mif_ready = 0;
    /* descriptors changed by plom_filter_update since the plan copied them */
    for (auto &ch : b->chains)
        for (size_t r = 0; r < ch.fs.size(); r++)
            if (ch.fs[r]->dirty)
                HIPCHK(hipMemcpyAsync(ch.ddf + r, ch.fs[r]->df, sizeof(plom_dfilter), hipMemcpyDeviceToDevice, c->stream));
    for (auto f : b->fs) f->dirty = 0;
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    if (use_graph) {
        if (!b->graph) {
            hipGraph_t g;
            /* the lanes' streams and the MIF side stream exist before the capture starts (none
             * is created inside it) */
            if (b->lanes > 1) {
                int rl = ctx_lanes(c, b->lanes);
                if (rl != PLOM_OK) return rl;
            } else {
                bool any_mif = false;
                for (auto &ch : b->chains) any_mif = any_mif || ch.any_mif > 0;
                mif_side tmp;
                if (any_mif && ctx_mif_side(c, &tmp) != PLOM_OK) return PLOM_ERR_HIP;
            }
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = batch_enqueue(c, b, 0, -1);
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc != PLOM_OK) return rc;
            if (e != hipSuccess) return set_err(PLOM_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
            HIPCHK(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(g));
        }
        c->kspans.clear();
        c->kspan_rounds = 0;
        HIPCHK(hipGraphLaunch(b->graph, c->stream));
    } else {
        int rc = batch_enqueue(c, b, c->ktiming, -1);
        if (rc != PLOM_OK) return rc;
    }
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    for (auto f : b->fs) f->has_run = 1;
    return PLOM_OK;
}

/* after a batch, each filter's own device descriptor gets the batch copy's state back
 * (plom_filter_read reads results through the filter's own buffers) */
static int batch_writeback(plom_ctx *c, plom_batch *b)
{
    for (auto &ch : b->chains)
        for (size_t r = 0; r < ch.fs.size(); r++)
            HIPCHK(hipMemcpyAsync(ch.fs[r]->df, ch.ddf + r, sizeof(plom_dfilter), hipMemcpyDeviceToDevice, c->stream));
    return PLOM_OK;
}

static int batch_free_chains(plom_ctx *c, plom_batch *b)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    if (b->graph) HIPCHK(hipGraphExecDestroy(b->graph));
    b->graph = nullptr;
    for (auto &ch : b->chains)
        if (ch.buf) HIPCHK(hipFree(ch.buf));
    b->chains.clear();
    return PLOM_OK;
}

static int batch_set_lanes(plom_ctx *c, plom_batch *b, int lanes)
{
    int rc = batch_free_chains(c, b);
    if (rc != PLOM_OK) return rc;
    if (lanes > b->nf) lanes = b->nf;
    b->lanes = lanes;
    b->chains.resize((size_t)lanes);
    for (int r = 0; r < b->nf; r++) b->chains[(size_t)(r % lanes)].fs.push_back(b->fs[(size_t)r]);
    for (auto &ch : b->chains) {
        rc = chain_build(c, &ch);
        if (rc != PLOM_OK) return rc;
        for (auto &g : ch.groups) {
            g.tpb = c->tpb;
            for (auto &t : b->tpb_rules)
                if (t.model == g.model && t.prec == g.prec && t.variant == g.variant && t.mif == g.mif) g.tpb = t.tpb;
        }
    }
    return PLOM_OK;
}

static int batch_destroy(plom_ctx *c, plom_batch *b)
{
    if (!b) return PLOM_OK;
    int rc = batch_free_chains(c, b);
    delete b;
    return rc;
}

extern "C" int plom_batch_create(plom_ctx *c, int nf, plom_filter **fs, plom_batch **out)
{
    if (!c || nf <= 0 || !fs || !out) return set_err(PLOM_ERR_ARG, "bad batch arguments");
    for (int i = 0; i < nf; i++)
        if (!fs[i]) return set_err(PLOM_ERR_ARG, "null filter %d", i);
    HIPCHK(hipSetDevice(c->device));
    plom_batch *b = new plom_batch();
    b->nf = nf;
    b->graph = nullptr;
    b->fs.assign(fs, fs + nf);
    std::stable_sort(b->fs.begin(), b->fs.end(), [](plom_filter *x, plom_filter *y) { return x->desc.N > y->desc.N; });
    int rc = batch_set_lanes(c, b, 1);
    if (rc != PLOM_OK) { (void)batch_destroy(c, b); return rc; }
    *out = b;
    return PLOM_OK;
}

extern "C" int plom_batch_configure(plom_ctx *c, plom_batch *b, int lanes)
{
    if (!c || !b) return set_err(PLOM_ERR_ARG, "null argument");
    if (lanes < 1 || lanes > PLOM_MAX_LANES) return set_err(PLOM_ERR_ARG, "lanes=%d outside [1, %d]", lanes, PLOM_MAX_LANES);
    HIPCHK(hipSetDevice(c->device));
    return batch_set_lanes(c, b, lanes);
}

/* time `rounds` launch rounds spread over the pass in the batch's current configuration
 * (a warm pass first: the first use of lanes and kernels) */
static int batch_time_rounds(plom_ctx *c, plom_batch *b, int rounds, double *ms)
{
    int rc = batch_enqueue(c, b, 0, rounds);
    if (rc != PLOM_OK) return rc;
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    rc = batch_enqueue(c, b, 0, rounds);
    if (rc != PLOM_OK) return rc;
    HIPCHK(hipEventRecord(c->ev[3], c->stream));
    HIPCHK(hipEventSynchronize(c->ev[3]));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->ev[2], c->ev[3]));
    *ms = t;
    return PLOM_OK;
}

/* The autotuner (the role of the reference's static J_chunk split of the particles over
 * threads, C/smc/main.c:298): first the number of lanes (1, 2, 4, 8) with the groups' current
 * tiles per block, then, lanes fixed, the tiles per block (1, 2, 4) of each (model,
 * precision, variant, MIF) group in turn, the others fixed; each candidate is the time of
 * `rounds` launch rounds spread over the pass.  Leaves the batch in the fastest
 * configuration found; results never depend on it. */
extern "C" int plom_batch_tune(plom_ctx *c, plom_batch *b, int rounds, plom_batch_tuning *out)
{
    if (!c || !b) return set_err(PLOM_ERR_ARG, "null argument");
    if (rounds < 1) return set_err(PLOM_ERR_ARG, "rounds=%d < 1", rounds);
    for (auto f : b->fs)
        if (f->desc.flags & PLOM_FLAG_MIF)
            return set_err(PLOM_ERR_ARG, "autotuning runs the filters and would consume the MIF starting clouds");
    HIPCHK(hipSetDevice(c->device));
    plom_batch_tuning res;
    memset(&res, 0, sizeof res);
    res.rounds = rounds;
    static const int lane_cand[4] = {1, 2, 4, 8};
    int best = 1;
    double best_ms = 0.0;
    for (int i = 0; i < 4; i++) {
        const int lanes = lane_cand[i];
        if (lanes > b->nf) continue;
        int rc = batch_set_lanes(c, b, lanes);
        if (rc == PLOM_OK) rc = batch_time_rounds(c, b, rounds, &res.lanes_ms[i]);
        if (rc != PLOM_OK) return rc;
        if (i == 0 || res.lanes_ms[i] < best_ms) { best = lanes; best_ms = res.lanes_ms[i]; }
    }
    res.lanes = best;
    /* the groups: the same keys in every chain */
    std::vector<tpb_rule> keys;
    for (auto f : b->fs) {
        const int mif = (f->desc.flags & PLOM_FLAG_MIF) != 0;
        bool seen = false;
        for (auto &k : keys)
            seen = seen || (k.model == f->desc.model_id && k.prec == f->desc.precision && k.variant == f->variant && k.mif == mif);
        if (!seen) keys.push_back(tpb_rule{f->desc.model_id, f->desc.precision, f->variant, mif, c->tpb});
    }
    static const int tpb_cand[3] = {1, 2, 4};
    for (size_t g = 0; g < keys.size(); g++) {
        tpb_rule &k = keys[g];
        plom_tune_group *tg = g < PLOM_TUNE_MAX_GROUPS ? &res.group[g] : nullptr;
        const bool multi = !k.mif && k.variant != 6;
        int tbest = k.tpb;
        double tbest_ms = -1.0;
        for (int i = 0; i < 3; i++) {
            if (!multi && tpb_cand[i] != 1) continue;
            k.tpb = tpb_cand[i];
            b->tpb_rules = keys;
            double ms = 0.0;
            int rc = batch_set_lanes(c, b, best);
            if (rc == PLOM_OK) rc = batch_time_rounds(c, b, rounds, &ms);
            if (rc != PLOM_OK) return rc;
            if (tg) tg->ms[i] = ms;
            if (tbest_ms < 0.0 || ms < tbest_ms) { tbest = k.tpb; tbest_ms = ms; }
        }
        k.tpb = tbest;
        if (tg) {
            tg->model_id = k.model; tg->precision = k.prec; tg->variant = k.variant; tg->mif = k.mif;
            tg->tiles_per_block = tbest;
        }
    }
    res.n_groups = (int32_t)(keys.size() < PLOM_TUNE_MAX_GROUPS ? keys.size() : PLOM_TUNE_MAX_GROUPS);
    b->tpb_rules = keys;
    int rc = batch_set_lanes(c, b, best);
    if (rc != PLOM_OK) return rc;
    /* the candidates' extra streams go; the batch's next launch creates what it uses */
    rc = ctx_release_lanes(c);
    if (rc != PLOM_OK) return rc;
    if (out) *out = res;
    return PLOM_OK;
}

extern "C" int plom_batch_autotune(plom_ctx *c, plom_batch *b, int rounds, int32_t *lanes_out, double *ms_out)
{
    plom_batch_tuning t;
    int rc = plom_batch_tune(c, b, rounds, &t);
    if (rc != PLOM_OK) return rc;
    if (lanes_out) *lanes_out = t.lanes;
    if (ms_out)
        for (int i = 0; i < 4; i++) ms_out[i] = t.lanes_ms[i];
    return PLOM_OK;
}

extern "C" int plom_ctx_set_tiles_per_block(plom_ctx *c, int tpb)
{
    if (!c) return set_err(PLOM_ERR_ARG, "null context");
    if (tpb != 1 && tpb != 2 && tpb != 4) return set_err(PLOM_ERR_ARG, "tiles per block %d is not 1, 2 or 4", tpb);
    c->tpb = tpb;
    return PLOM_OK;
}

extern "C" int plom_batch_launch(plom_ctx *c, plom_batch *b, int use_graph)
{
    if (!c || !b) return set_err(PLOM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    int rc = batch_launch(c, b, use_graph);
    if (rc == PLOM_OK) rc = batch_writeback(c, b);
    return rc;
}

extern "C" int plom_batch_destroy(plom_ctx *c, plom_batch *b)
{
    if (!c) return set_err(PLOM_ERR_ARG, "null context");
    return batch_destroy(c, b);
}

extern "C" int plom_batch_run(plom_ctx *c, int nf, plom_filter **fs, int use_graph)
{
    if (!c || nf < 0 || (nf > 0 && !fs)) return set_err(PLOM_ERR_ARG, "bad batch arguments");
    if (nf == 0) return PLOM_OK;
    plom_batch *b = nullptr;
    int rc = plom_batch_create(c, nf, fs, &b);
    if (rc != PLOM_OK) return rc;
    rc = plom_batch_launch(c, b, use_graph);
    int rc2 = batch_destroy(c, b);
    return rc != PLOM_OK ? rc : rc2;
}

/* ------------------------------------------------------------------ server */
/* Continuous batching: requests wait in a FIFO on the host; at every step the server
 * admits the queue head while its device bytes fit the HBM budget (and the slot limit),
 * then advances the requests of each lane by one observation interval with one batched
 * launch round (one propagate launch per (model, precision, variant) group, then the
 * weight / scan / advance launches over all of them), and retires the requests whose last
 * interval was just enqueued (a HIP event marks their completion).  A request's result
 * depends only on its descriptor and inputs (its random key is its own stream id), never
 * on the mix around it.
 *
 * Two lanes: requests whose work per interval (J x size_proj) reaches the heavy threshold
 * run in lane 1, on a stream of their own, so that the short intervals of the small
 * requests (lane 0, the context stream) do not wait behind theirs; the two lanes run
 * concurrently on the device.  While small requests are in flight, the heavy lane's
 * propagate launches cover at most heavy_blocks blocks each (one block per CU by default):
 * a launch over all of a heavy request's tiles would hold every CU's slots, and each small
 * launch would wait for heavy blocks to retire before its own could start.  Each lane keeps
 * at most max_inflight_steps steps queued;
 * a step skips a lane whose window is full and blocks only when no lane can advance.
 * Host-to-device uploads come from a pool of page-locked blocks, released once the lane
 * step that used them is known to be complete. */
#define SRV_QUEUED 0
#define SRV_RUNNING 1
#define SRV_DONE 2
#define SRV_LANES 2

/* page-locked host blocks in power-of-two size classes, reused, freed at the end */
struct pinned_pool {
    std::vector<std::vector<char *>> free_by_class;
    std::vector<std::pair<char *, int>> all;
    static int size_class(size_t bytes)
    {
        int k = 12;
        while (((size_t)1 << k) < bytes) k++;
        return k;
    }
    int get(size_t bytes, char **out, int *cls)
    {
        const int k = size_class(bytes);
        if ((int)free_by_class.size() <= k) free_by_class.resize((size_t)k + 1);
        if (!free_by_class[(size_t)k].empty()) {
            *out = free_by_class[(size_t)k].back();
            free_by_class[(size_t)k].pop_back();
        } else {
            hipError_t e = hipHostMalloc((void **)out, (size_t)1 << k, hipHostMallocDefault);
            if (e != hipSuccess) return set_err(PLOM_ERR_NOMEM, "hipHostMalloc(%zu): %s", (size_t)1 << k, hipGetErrorString(e));
            all.push_back({*out, k});
        }
        *cls = k;
        return PLOM_OK;
    }
    void put(char *p, int k) { free_by_class[(size_t)k].push_back(p); }
    void release_all()
    {
        for (auto &pk : all) (void)hipHostFree(pk.first);
        all.clear();
        free_by_class.clear();
    }
};

struct srv_request {
    uint64_t id;
    plom_filter_desc desc;
    std::vector<double> times, data, pv, x0;
    /* time-varying covariates and the drift map, copied at submission */
    std::vector<plom_covariate> cov;
    std::vector<double> cov_x, cov_y;
    std::vector<int32_t> dm_i;      /* group, tr, scale, duration, vol_slot */
    std::vector<double> dm_d;       /* min, max, multiplier */
    plom_drift_map dmap = {};
    int has_drift = 0;
    size_t cov_vals = 0, spl_vals = 0;  /* sizes of the covariate table and splines in the layout */
    plom_filter *f = nullptr;
    int slot = -1;
    int lane = 0;
    int progress = 0;          /* intervals enqueued */
    int state = SRV_QUEUED;
    hipEvent_t done_ev = nullptr;
    int64_t bytes = 0;
};

struct srv_lane {
    hipStream_t stream = nullptr;
    std::vector<srv_request *> active;
    int2 *d_tiles = nullptr;                   /* tile maps: all tiles, then per group */
    size_t tiles_cap = 0;
    int *d_slots_active = nullptr;             /* slot index of every active request */
    int dirty = 1;                             /* active set changed: rebuild the maps */
    struct group { int model, prec, variant; size_t off, n; };
    std::vector<group> groups;
    size_t n_all = 0;
    int64_t steps = 0;
    std::vector<hipEvent_t> step_ev;           /* ring of max_inflight_steps events */
    std::vector<std::vector<std::pair<char *, int>>> step_release;  /* pinned blocks per ring slot */
};

struct plom_server {
    plom_ctx *c;
    int64_t budget, in_use, heavy;
    int heavy_blocks;                          /* blocks per heavy propagate launch while small
                                                  requests are in flight (0: no limit) */
    int max_active;
    uint64_t next_id;
    std::vector<srv_request *> queue;          /* FIFO */
    std::vector<srv_request *> finished;       /* waiting for their event */
    std::vector<srv_request *> done;           /* complete, result not taken yet */
    int capacity;                              /* descriptor slots on the device */
    plom_dfilter *d_slots;
    std::vector<int> free_slots;
    int depth;                                 /* max steps in flight per lane */
    srv_lane lanes[SRV_LANES];
    pinned_pool pool;
};

static int srv_n_active(const plom_server *S)
{
    int n = 0;
    for (int l = 0; l < SRV_LANES; l++) n += (int)S->lanes[l].active.size();
    return n;
}

/* the lane may start its next step: its ring slot is free (the step that used it last
 * has completed); then the slot's pinned blocks go back to the pool */
static int srv_lane_ready(plom_server *S, srv_lane *L, int block)
{
    const size_t k = (size_t)(L->steps % S->depth);
    if (L->steps >= S->depth) {
        hipEvent_t e = L->step_ev[k];
        if (block) {
            HIPCHK(hipEventSynchronize(e));
        } else {
            hipError_t q = hipEventQuery(e);
            if (q == hipErrorNotReady) { (void)hipGetLastError(); return 0; }
            if (q != hipSuccess) return set_err(PLOM_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(q));
        }
    }
    for (auto &pk : L->step_release[k]) S->pool.put(pk.first, pk.second);
    L->step_release[k].clear();
    return 1;
}

static int srv_stage(plom_server *S, srv_lane *L, const void *src, size_t bytes, void *dst)
{
    char *p;
    int k;
    int rc = S->pool.get(bytes, &p, &k);
    if (rc != PLOM_OK) return rc;
    memcpy(p, src, bytes);
    L->step_release[(size_t)(L->steps % S->depth)].push_back({p, k});
    HIPCHK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, L->stream));
    return PLOM_OK;
}

static int srv_rebuild_maps(plom_server *S, srv_lane *L)
{
    std::vector<int2> h;
    L->groups.clear();
    for (auto r : L->active)
        for (int t = 0; t < r->f->h.T; t++) h.push_back(make_int2(r->slot, t));
    L->n_all = h.size();
    for (auto r : L->active) {
        bool seen = false;
        for (auto &gg : L->groups)
            if (gg.model == r->desc.model_id && gg.prec == r->desc.precision && gg.variant == r->f->variant) seen = true;
        if (!seen) L->groups.push_back(srv_lane::group{r->desc.model_id, r->desc.precision, r->f->variant, 0, 0});
    }
    for (auto &g : L->groups) {
        g.off = h.size();
        for (auto r : L->active)
            if (g.model == r->desc.model_id && g.prec == r->desc.precision && g.variant == r->f->variant)
                for (int t = 0; t < r->f->h.T; t++) h.push_back(make_int2(r->slot, t));
        g.n = h.size() - g.off;
    }
    if (h.size() > L->tiles_cap) {
        if (L->d_tiles) HIPCHK(hipFreeAsync(L->d_tiles, L->stream));
        L->tiles_cap = h.size() * 2 + 1024;
        HIPCHK(hipMallocAsync((void **)&L->d_tiles, L->tiles_cap * sizeof(int2), L->stream));
    }
    if (!h.empty()) {
        int rc = srv_stage(S, L, h.data(), h.size() * sizeof(int2), L->d_tiles);
        if (rc != PLOM_OK) return rc;
    }
    std::vector<int> ha;
    for (auto r : L->active) ha.push_back(r->slot);
    if (!ha.empty()) {
        int rc = srv_stage(S, L, ha.data(), ha.size() * sizeof(int), L->d_slots_active);
        if (rc != PLOM_OK) return rc;
    }
    L->dirty = 0;
    return PLOM_OK;
}

static void srv_free(plom_server *S)
{
    for (int l = 0; l < SRV_LANES; l++) {
        srv_lane *L = &S->lanes[l];
        for (auto e : L->step_ev) (void)hipEventDestroy(e);
        if (L->d_slots_active) (void)hipFree(L->d_slots_active);
        if (l > 0 && L->stream) (void)hipStreamDestroy(L->stream);
    }
    if (S->d_slots) (void)hipFree(S->d_slots);
    S->pool.release_all();
    delete S;
}

extern "C" int plom_server_create(plom_ctx *c, const plom_server_config *cfg, plom_server **out)
{
    if (!c || !out) return set_err(PLOM_ERR_ARG, "null argument");
    if (cfg && (cfg->max_active < 0 || cfg->max_inflight_steps < 0)) return set_err(PLOM_ERR_ARG, "negative server limit");
    HIPCHK(hipSetDevice(c->device));
    plom_server *S = new (std::nothrow) plom_server();
    if (!S) return set_err(PLOM_ERR_NOMEM, "host allocation failed");
    S->c = c;
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    S->budget = (cfg && cfg->hbm_budget_bytes > 0) ? cfg->hbm_budget_bytes : (int64_t)(0.8 * (double)free_b);
    S->max_active = (cfg && cfg->max_active > 0) ? cfg->max_active : 4096;
    S->depth = (cfg && cfg->max_inflight_steps > 0) ? cfg->max_inflight_steps : 4;
    S->heavy = (cfg && cfg->heavy_threshold != 0) ? cfg->heavy_threshold : (int64_t)4000000;
    if (cfg && cfg->heavy_blocks < 0) {
        S->heavy_blocks = 0;
    } else if (cfg && cfg->heavy_blocks > 0) {
        S->heavy_blocks = cfg->heavy_blocks;
    } else {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || ncu <= 0) {
            srv_free(S);
            return set_err(PLOM_ERR_HIP, "the device's CU count is unavailable");
        }
        S->heavy_blocks = ncu;
    }
    S->capacity = S->max_active;
    S->in_use = 0;
    S->next_id = 1;
    S->d_slots = nullptr;
    hipError_t e = hipMalloc((void **)&S->d_slots, (size_t)S->capacity * sizeof(plom_dfilter));
    if (e != hipSuccess) { srv_free(S); return set_err(PLOM_ERR_NOMEM, "hipMalloc slots: %s", hipGetErrorString(e)); }
    /* the batch lanes' streams go first, so that the heavy lane's stream gets a hardware
     * queue of its own instead of sharing the context stream's */
    int rl = ctx_release_lanes(c);
    if (rl != PLOM_OK) { srv_free(S); return rl; }
    for (int l = 0; l < SRV_LANES; l++) {
        srv_lane *L = &S->lanes[l];
        if (l == 0) {
            L->stream = c->stream;
        } else {
            e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
            if (e != hipSuccess) { L->stream = nullptr; srv_free(S); return set_err(PLOM_ERR_HIP, "heavy lane stream: %s", hipGetErrorString(e)); }
        }
        e = hipMalloc((void **)&L->d_slots_active, (size_t)S->capacity * sizeof(int));
        if (e != hipSuccess) { L->d_slots_active = nullptr; srv_free(S); return set_err(PLOM_ERR_NOMEM, "hipMalloc slots: %s", hipGetErrorString(e)); }
        L->step_release.resize((size_t)S->depth);
        for (int i = 0; i < S->depth; i++) {
            hipEvent_t ev;
            e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (e != hipSuccess) { srv_free(S); return set_err(PLOM_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e)); }
            L->step_ev.push_back(ev);
        }
    }
    for (int i = S->capacity - 1; i >= 0; i--) S->free_slots.push_back(i);
    *out = S;
    return PLOM_OK;
}

extern "C" int plom_server_submit(plom_server *S, const plom_filter_desc *d, const plom_filter_inputs *in, uint64_t *req_id)
{
    if (!S || !d || !in || !req_id) return set_err(PLOM_ERR_ARG, "null argument");
    plom_model_info mi;
    int rc = check_desc(d, &mi);
    if (rc != PLOM_OK) return rc;
    if (d->flags & (PLOM_FLAG_STORE_ANCESTRY | PLOM_FLAG_FINAL_STATE | PLOM_FLAG_HAT | PLOM_FLAG_MIF | PLOM_FLAG_ALL_STATES))
        return set_err(PLOM_ERR_ARG, "served requests return log-likelihoods and lineages (no ancestry, final state, hat or MIF)");
    if (!in->times || !in->pv || !in->x0 || (d->N > 0 && !in->data)) return set_err(PLOM_ERR_ARG, "null input array");
    for (int n = 0; n < d->N; n++)
        if (!(in->times[n + 1] > in->times[n])) return set_err(PLOM_ERR_ARG, "times must increase (n=%d)", n);
    /* the covariates and the drift map are checked now, so that admission cannot fail on them */
    input_prep prep;
    rc = prepare_inputs(d, mi, in, &prep);
    if (rc != PLOM_OK) return rc;
    layout L;
    make_layout(d, &mi, &L, prep.covtab.size(), 0, 0, prep.spl_val.size());
    if ((int64_t)L.total > S->budget)
        return set_err(PLOM_ERR_NOMEM, "request needs %zu device bytes, more than the server budget %lld", L.total, (long long)S->budget);
    srv_request *r = new (std::nothrow) srv_request();
    if (!r) return set_err(PLOM_ERR_NOMEM, "host allocation failed");
    r->id = S->next_id++;
    r->desc = *d;
    r->times.assign(in->times, in->times + d->N + 1);
    if (d->N > 0) r->data.assign(in->data, in->data + (size_t)d->N * mi.n_ts);
    r->pv.assign(in->pv, in->pv + mi.pv_size);
    r->x0.assign(in->x0, in->x0 + mi.size_proj);
    r->cov_vals = prep.covtab.size();
    r->spl_vals = prep.spl_val.size();
    if (prep.cov_on) {
        for (int i = 0; i < mi.n_fix; i++) {
            r->cov_x.insert(r->cov_x.end(), in->cov[i].x, in->cov[i].x + in->cov[i].n);
            r->cov_y.insert(r->cov_y.end(), in->cov[i].y, in->cov[i].y + in->cov[i].n);
        }
        r->cov.resize((size_t)mi.n_fix);
        size_t o = 0;
        for (int i = 0; i < mi.n_fix; i++) {
            r->cov[(size_t)i] = {in->cov[i].n, 0, r->cov_x.data() + o, r->cov_y.data() + o};
            o += (size_t)in->cov[i].n;
        }
    }
    if (prep.drift_on) {
        const size_t ND = (size_t)mi.n_drift * (size_t)mi.n_cac;
        const plom_drift_map *m = in->drift;
        for (const int32_t *a : {m->group, m->tr, m->scale, m->duration, m->vol_slot}) r->dm_i.insert(r->dm_i.end(), a, a + ND);
        for (const double *a : {m->min, m->max, m->multiplier}) r->dm_d.insert(r->dm_d.end(), a, a + ND);
        r->dmap.group = r->dm_i.data();
        r->dmap.tr = r->dm_i.data() + ND;
        r->dmap.scale = r->dm_i.data() + 2 * ND;
        r->dmap.duration = r->dm_i.data() + 3 * ND;
        r->dmap.vol_slot = r->dm_i.data() + 4 * ND;
        r->dmap.min = r->dm_d.data();
        r->dmap.max = r->dm_d.data() + ND;
        r->dmap.multiplier = r->dm_d.data() + 2 * ND;
        r->has_drift = 1;
    }
    r->bytes = (int64_t)L.total;
    r->lane = (S->heavy > 0 && (int64_t)d->J * mi.size_proj >= S->heavy) ? 1 : 0;
    S->queue.push_back(r);
    *req_id = r->id;
    return PLOM_OK;
}

/* move the finished requests whose completion event has passed to `done` */
static int srv_poll(plom_server *S)
{
    for (size_t i = 0; i < S->finished.size();) {
        srv_request *r = S->finished[i];
        hipError_t q = hipEventQuery(r->done_ev);
        if (q == hipSuccess) {
            r->state = SRV_DONE;
            S->free_slots.push_back(r->slot);
            S->done.push_back(r);
            S->finished.erase(S->finished.begin() + (long)i);
        } else if (q == hipErrorNotReady) {
            (void)hipGetLastError();
            i++;
        } else {
            return set_err(PLOM_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(q));
        }
    }
    return PLOM_OK;
}

/* small requests are active or queued */
static bool srv_light_busy(const plom_server *S)
{
    if (!S->lanes[0].active.empty()) return true;
    for (auto r : S->queue)
        if (r->lane == 0) return true;
    return false;
}

/* one interval for every active request of a lane whose ring has room.  The heavy lane,
 * while small requests are in flight, covers each group's tiles in launches of at most
 * heavy_blocks blocks (consecutive tile ranges; the stream runs them one after another, and
 * the tiles are independent, so the results are those of one launch). */
static int srv_lane_step(plom_server *S, srv_lane *L)
{
    plom_ctx *c = S->c;
    if (L->dirty) {
        int rc = srv_rebuild_maps(S, L);
        if (rc != PLOM_OK) return rc;
    }
    (void)hipGetLastError();
    const bool limit = L == &S->lanes[1] && S->heavy_blocks > 0 && srv_light_busy(S);
    for (auto &g : L->groups) {
        const size_t chunk = limit ? (size_t)S->heavy_blocks * (size_t)c->tpb : g.n;
        for (size_t o = 0; o < g.n; o += chunk)
            model_at(g.model)->launch_prop(g.prec, g.variant, 0, (int)std::min(chunk, g.n - o), c->tpb, L->stream,
                                           S->d_slots, L->d_tiles + g.off + o);
    }
    enqueue_resample(L->stream, S->d_slots, (const int2 *)L->d_tiles, (int)L->n_all, L->d_slots_active, (int)L->active.size());
    HIPCHK(hipGetLastError());
    for (size_t i = 0; i < L->active.size();) {
        srv_request *r = L->active[i];
        r->progress++;
        if (r->progress >= r->desc.N) {
            /* the slot descriptor carries the interval counter: copy it back, mark done */
            if (r->desc.N > 0 && (r->desc.flags & PLOM_FLAG_LINEAGE))
                model_at(r->desc.model_id)->launch_lineage(r->desc.precision, r->f->variant, L->stream, S->d_slots + r->slot);
            HIPCHK(hipMemcpyAsync(r->f->df, S->d_slots + r->slot, sizeof(plom_dfilter), hipMemcpyDeviceToDevice, L->stream));
            HIPCHK(hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming));
            HIPCHK(hipEventRecord(r->done_ev, L->stream));
            r->f->has_run = 1;
            /* the slot goes back to the free list only when done_ev has passed (srv_poll):
             * the lane's stream still has this request's last steps, lineage and copy-back
             * queued, and they read the slot; a request admitted into it meanwhile on the
             * other lane's stream would overwrite the slot under them */
            S->finished.push_back(r);
            L->active.erase(L->active.begin() + (long)i);
            L->dirty = 1;
        } else {
            i++;
        }
    }
    HIPCHK(hipEventRecord(L->step_ev[(size_t)(L->steps % S->depth)], L->stream));
    L->steps++;
    (void)c;
    return PLOM_OK;
}

/* admit, advance every lane that can start a step by one interval, retire; returns counts */
extern "C" int plom_server_step(plom_server *S, int32_t *n_active, int32_t *n_queued, int32_t *n_done)
{
    if (!S) return set_err(PLOM_ERR_ARG, "null server");
    plom_ctx *c = S->c;
    HIPCHK(hipSetDevice(c->device));
    int rc = srv_poll(S);
    if (rc != PLOM_OK) return rc;
    /* which lanes can take a step now (a lane with queued work counts as busy) */
    int ready[SRV_LANES];
    for (int l = 0; l < SRV_LANES; l++) {
        ready[l] = 0;
        srv_lane *L = &S->lanes[l];
        bool wants = !L->active.empty();
        for (auto r : S->queue) wants = wants || r->lane == l;
        if (!wants) continue;
        rc = srv_lane_ready(S, L, 0);
        if (rc < 0) return rc;
        ready[l] = rc;
    }
    if (!ready[0] && !ready[1]) {
        /* nothing can start: wait for the lane that will free a slot first (the light one
         * if it is busy) */
        for (int l = 0; l < SRV_LANES; l++) {
            srv_lane *L = &S->lanes[l];
            bool wants = !L->active.empty();
            for (auto r : S->queue) wants = wants || r->lane == l;
            if (!wants) continue;
            rc = srv_lane_ready(S, L, 1);
            if (rc < 0) return rc;
            ready[l] = 1;
            break;
        }
    }
    /* admission by free HBM, FIFO: the queue head blocks the queue */
    while (!S->queue.empty() && srv_n_active(S) < S->max_active && !S->free_slots.empty()) {
        srv_request *r = S->queue.front();
        if (S->in_use + r->bytes > S->budget || !ready[r->lane]) break;
        srv_lane *L = &S->lanes[r->lane];
        plom_model_info mi = model_at(r->desc.model_id)->info;
        layout Lo;
        make_layout(&r->desc, &mi, &Lo, r->cov_vals, 0, 0, r->spl_vals);
        char *img;
        int k;
        rc = S->pool.get(image_bytes(&Lo), &img, &k);
        if (rc != PLOM_OK) return rc;
        L->step_release[(size_t)(L->steps % S->depth)].push_back({img, k});
        plom_filter_inputs in = {r->times.data(), r->data.data(), r->pv.data(), r->x0.data(), nullptr,
                                 r->cov.empty() ? nullptr : r->cov.data(), r->has_drift ? &r->dmap : nullptr};
        rc = filter_create_impl(c, &r->desc, &in, img, &r->f, L->stream);
        if (rc != PLOM_OK) return rc;
        r->slot = S->free_slots.back();
        S->free_slots.pop_back();
        /* the slot descriptor starts as the filter's own (uploaded just before) */
        HIPCHK(hipMemcpyAsync(S->d_slots + r->slot, r->f->df, sizeof(plom_dfilter), hipMemcpyDeviceToDevice, L->stream));
        S->in_use += r->bytes;
        r->state = SRV_RUNNING;
        L->active.push_back(r);
        S->queue.erase(S->queue.begin());
        L->dirty = 1;
    }
    for (int l = 0; l < SRV_LANES; l++) {
        srv_lane *L = &S->lanes[l];
        if (!ready[l] || L->active.empty()) continue;
        rc = srv_lane_step(S, L);
        if (rc != PLOM_OK) return rc;
    }
    if (n_active) *n_active = (int32_t)srv_n_active(S);
    if (n_queued) *n_queued = (int32_t)S->queue.size();
    if (n_done) *n_done = (int32_t)S->done.size();
    return PLOM_OK;
}

static srv_request *srv_find(plom_server *S, uint64_t id, std::vector<srv_request *> **where)
{
    std::vector<srv_request *> *lists[] = {&S->queue, &S->lanes[0].active, &S->lanes[1].active, &S->finished, &S->done};
    for (auto L : lists)
        for (auto r : *L)
            if (r->id == id) { if (where) *where = L; return r; }
    return nullptr;
}

extern "C" int plom_server_status(plom_server *S, uint64_t id)
{
    if (!S) return set_err(PLOM_ERR_ARG, "null server");
    int rc = srv_poll(S);
    if (rc != PLOM_OK) return rc;
    srv_request *r = srv_find(S, id, nullptr);
    if (!r) return set_err(PLOM_ERR_ARG, "unknown request %llu", (unsigned long long)id);
    return r->state;
}

extern "C" int plom_server_done(plom_server *S, uint64_t *ids, int32_t cap, int32_t *n)
{
    if (!S || !n || cap < 0 || (cap > 0 && !ids)) return set_err(PLOM_ERR_ARG, "bad arguments");
    int rc = srv_poll(S);
    if (rc != PLOM_OK) return rc;
    int32_t k = 0;
    for (auto r : S->done) {
        if (k >= cap) break;
        ids[k++] = r->id;
    }
    *n = k;
    return PLOM_OK;
}

extern "C" int plom_server_result(plom_server *S, uint64_t id, plom_filter_outputs *o)
{
    if (!S || !o) return set_err(PLOM_ERR_ARG, "null argument");
    if (o->ancestry || o->final_state || o->final_logw || o->hat || o->pred_res)
        return set_err(PLOM_ERR_ARG, "served requests return log-likelihoods and lineages (no ancestry, final state or hat)");
    std::vector<srv_request *> *where = nullptr;
    srv_request *r = srv_find(S, id, &where);
    if (!r) return set_err(PLOM_ERR_ARG, "unknown request %llu", (unsigned long long)id);
    if (!r->done_ev) return set_err(PLOM_ERR_STATE, "request %llu has not finished", (unsigned long long)id);
    HIPCHK(hipSetDevice(S->c->device));
    HIPCHK(hipEventSynchronize(r->done_ev));
    if (where == &S->finished) {
        /* taken before a poll moved it to `done`: its completion event has now passed, so
         * its slot goes back to the free list here (srv_poll will never see it again) */
        r->state = SRV_DONE;
        S->free_slots.push_back(r->slot);
        where->erase(std::find(where->begin(), where->end(), r));
        S->done.push_back(r);
        where = &S->done;
    }
    /* the request's work is complete: read on the side stream, past the queued steps */
    int rc = filter_read_on(S->c, r->f, o, S->c->io);
    if (rc != PLOM_OK) return rc;
    HIPCHK(hipEventDestroy(r->done_ev));
    HIPCHK(hipFreeAsync(r->f->dev, S->lanes[r->lane].stream));
    S->in_use -= r->bytes; /* the request's device memory counts until its result is taken */
    delete r->f;
    where->erase(std::find(where->begin(), where->end(), r));
    delete r;
    return PLOM_OK;
}

extern "C" int plom_server_destroy(plom_server *S)
{
    if (!S) return PLOM_OK;
    plom_ctx *c = S->c;
    for (int l = 0; l < SRV_LANES; l++)
        if (S->lanes[l].stream) HIPCHK(hipStreamSynchronize(S->lanes[l].stream));
    std::vector<srv_request *> *lists[] = {&S->queue, &S->lanes[0].active, &S->lanes[1].active, &S->finished, &S->done};
    for (auto L : lists)
        for (auto r : *L) {
            if (r->done_ev) (void)hipEventDestroy(r->done_ev);
            if (r->f) { (void)hipFreeAsync(r->f->dev, S->lanes[r->lane].stream); delete r->f; }
            delete r;
        }
    for (int l = 0; l < SRV_LANES; l++) {
        srv_lane *L = &S->lanes[l];
        if (L->d_tiles) HIPCHK(hipFreeAsync(L->d_tiles, L->stream));
        HIPCHK(hipStreamSynchronize(L->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    srv_free(S);
    return PLOM_OK;
}

extern "C" int plom_ctx_set_kernel_timing(plom_ctx *c, int on)
{
    if (!c) return set_err(PLOM_ERR_ARG, "null context");
    c->ktiming = on ? 1 : 0;
    return PLOM_OK;
}

extern "C" int plom_last_timing(plom_ctx *c, double *total_ms, double *prop_ms, double *scan_ms, int32_t *n_intervals,
                                int32_t *n_prop_launches)
{
    if (!c) return set_err(PLOM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->last_total_ms = ms;
    double p = 0.0, q = 0.0;
    for (const auto &sp : c->kspans) {
        float a = 0.f;
        HIPCHK(hipEventElapsedTime(&a, c->kev[(size_t)sp.a], c->kev[(size_t)sp.b]));
        (sp.kind == 0 ? p : q) += a;
    }
    const int timed = !c->kspans.empty();
    if (total_ms) *total_ms = ms;
    if (prop_ms) *prop_ms = timed ? p : -1.0;
    if (scan_ms) *scan_ms = timed ? q : -1.0;
    if (n_intervals) *n_intervals = timed ? c->kspan_rounds : 0;
    if (n_prop_launches) *n_prop_launches = timed ? c->kprop_launches : 0;
    return PLOM_OK;
}

extern "C" int plom_last_timing_lanes(plom_ctx *c, double *prop_ms, double *scan_ms, int32_t cap, int32_t *n_lanes)
{
    if (!c || !n_lanes || cap < 0 || (cap > 0 && (!prop_ms || !scan_ms))) return set_err(PLOM_ERR_ARG, "bad arguments");
    HIPCHK(hipStreamSynchronize(c->stream));
    int nl = 0;
    for (const auto &sp : c->kspans) nl = sp.lane + 1 > nl ? sp.lane + 1 : nl;
    for (int l = 0; l < cap; l++) { prop_ms[l] = 0.0; scan_ms[l] = 0.0; }
    for (const auto &sp : c->kspans) {
        if (sp.lane >= cap) continue;
        float a = 0.f;
        HIPCHK(hipEventElapsedTime(&a, c->kev[(size_t)sp.a], c->kev[(size_t)sp.b]));
        (sp.kind == 0 ? prop_ms : scan_ms)[sp.lane] += a;
    }
    *n_lanes = nl;
    return PLOM_OK;
}

extern "C" int plom_last_timing_busy(plom_ctx *c, double *prop_busy_ms, double *scan_busy_ms)
{
    if (!c || !prop_busy_ms || !scan_busy_ms) return set_err(PLOM_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    *prop_busy_ms = -1.0;
    *scan_busy_ms = -1.0;
    if (c->kspans.empty()) return PLOM_OK;
    /* each span as [start, end] from the run's first event (recorded on the context stream
     * before the lanes fork from it), then the length of the union per kind */
    std::vector<std::pair<float, float>> iv[2];
    for (const auto &sp : c->kspans) {
        float a = 0.f, b = 0.f;
        HIPCHK(hipEventElapsedTime(&a, c->ev[0], c->kev[(size_t)sp.a]));
        HIPCHK(hipEventElapsedTime(&b, c->ev[0], c->kev[(size_t)sp.b]));
        iv[sp.kind == 0 ? 0 : 1].push_back({a, b});
    }
    double out[2] = { 0.0, 0.0 };
    for (int kd = 0; kd < 2; kd++) {
        auto &v = iv[kd];
        std::sort(v.begin(), v.end());
        double lo = 0.0, hi = 0.0, sum = 0.0;
        bool open = false;
        for (const auto &p : v) {
            if (open && p.first <= hi) { if (p.second > hi) hi = p.second; continue; }
            if (open) sum += hi - lo;
            lo = p.first; hi = p.second; open = true;
        }
        if (open) sum += hi - lo;
        out[kd] = sum;
    }
    *prop_busy_ms = out[0];
    *scan_busy_ms = out[1];
    return PLOM_OK;
}

extern "C" int plom_selftest_math(plom_ctx *c, int n, const double *in, double *out)
{
    if (!c || n < 0 || (n > 0 && (!in || !out))) return set_err(PLOM_ERR_ARG, "bad arguments");
    if (n == 0) return PLOM_OK;
    HIPCHK(hipSetDevice(c->device));
    double *d = nullptr;
    HIPCHK(hipMalloc((void **)&d, (size_t)n * 9 * 8));   /* n inputs, then 8 results each */
    HIPCHK(hipMemcpy(d, in, (size_t)n * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_selftest_math, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, d, d + n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(out, d + n, (size_t)n * 8 * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipFree(d));
    return PLOM_OK;
}

extern "C" int plom_selftest_copy(plom_ctx *c, int64_t nbytes, int reps, double *ms_per_copy)
{
    if (!c || nbytes < 16 || reps < 1) return set_err(PLOM_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(c->device));
    const size_t n = (size_t)nbytes / 16;
    double2 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc((void **)&a, n * 16));
    HIPCHK(hipMalloc((void **)&b, n * 16));
    HIPCHK(hipMemsetAsync(a, 0, n * 16, c->stream));
    const int blocks = 256 * 8 * 4;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, c->stream, (const double2 *)a, b, n);
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, c->stream, (const double2 *)a, b, n);
    HIPCHK(hipEventRecord(c->ev[3], c->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    if (ms_per_copy) *ms_per_copy = ms / reps;
    HIPCHK(hipFree(a));
    HIPCHK(hipFree(b));
    return PLOM_OK;
}

static int selftest_normals(plom_ctx *c, uint32_t seed, uint32_t stream, uint32_t n0, uint32_t j, int n, int fp32,
                            double *out)
{
    if (!c || n < 0 || (n > 0 && !out)) return set_err(PLOM_ERR_ARG, "bad arguments");
    if (n == 0) return PLOM_OK;
    HIPCHK(hipSetDevice(c->device));
    double *d = nullptr;
    HIPCHK(hipMalloc((void **)&d, (size_t)n * 8));
    hipLaunchKernelGGL(k_selftest_normals, dim3(1), dim3(64), 0, c->stream, seed, stream, n0, j, n, fp32, d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(out, d, (size_t)n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipFree(d));
    return PLOM_OK;
}

extern "C" int plom_selftest_normals(plom_ctx *c, uint32_t seed, uint32_t stream, uint32_t n0, uint32_t j, int n, double *out)
{
    return selftest_normals(c, seed, stream, n0, j, n, 0, out);
}

extern "C" int plom_selftest_normals32(plom_ctx *c, uint32_t seed, uint32_t stream, uint32_t n0, uint32_t j, int n,
                                       double *out)
{
    return selftest_normals(c, seed, stream, n0, j, n, 1, out);
}
