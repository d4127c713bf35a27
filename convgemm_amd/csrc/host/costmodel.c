/*
 * costmodel.c -- learned cost model for tile-configuration selection.
 *
 * An MLP  F=33 features -> 64 -> 64 -> 1  (ReLU) predicts the standardised
 * log time of one (problem, candidate) pair.  Features (cg_features), log2
 * where a size:
 *   0-5   operation one-hot (forward / weight gradient / backward data x NHWC, NCHW)
 *   6-8   operand type one-hot (FP32, BF16, FP8);  9  16-byte loads (vec)
 *   10-14 pixels P, channels C, kn, taps, stride product
 *   15-17 GEMM rows M, columns N, reduction length (K tiles x BK)
 *   18-21 tile BM, BN, bytes per K step of a row, split-K factor
 *   22    workgroups launched;  23  tile efficiency (useful / computed outputs)
 *   24    waves of workgroups over 256 CUs at the tile's LDS occupancy
 *   25    computed FLOPs (with tile padding)
 *   26    pipelined kernel;  27  stream-K;  28  waves per workgroup;  29  LDS stages
 *   30-31 raster order one-hot (order 0, the N panels of an M panel adjacent, is neither)
 *   32    the halo-tile kernel
 * It is trained on every timing
 * the autotuner recorded (tune.c) by the device kernels of mlp.hip --
 * forward, backward and Adam on the GPU -- and evaluated by the same forward
 * kernels to rank the candidates of a problem that was never tuned.  The
 * analytic planner (cg_default_plan) is the "no learning" baseline it is
 * measured against.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "convgemm.h"
#include "convgemm_tune.h"
#include "../cg_internal.h"
#include "cg_host.h"

#define CM_F 33
#define CM_F_V1 30 /* models saved with fewer features (before the raster-order and halo
                      * features) load with the later features' weights at 0 */
#define CM_H 64
#define CM_NP (CM_F * CM_H + CM_H + CM_H * CM_H + CM_H + CM_H + 1)
#define OFF_W1 0
#define OFF_B1 (CM_F * CM_H)
#define OFF_W2 (OFF_B1 + CM_H)
#define OFF_B2 (OFF_W2 + CM_H * CM_H)
#define OFF_W3 (OFF_B2 + CM_H)
#define OFF_B3 (OFF_W3 + CM_H)

int cg_mlp_dense_fwd(const float *A, const float *W, const float *b, float *H, int N, int I, int J, int relu,
                     void *stream);
int cg_mlp_mse_grad(const float *Y, const float *T, float *dY, float *partial, int N, void *stream);
int cg_mlp_dense_bwd_act(const float *dZ, const float *W, const float *A, float *dA, int N, int I, int J,
                         void *stream);
int cg_mlp_dense_wgrad(const float *A, const float *dZ, float *dW, float *db, int N, int I, int J, void *stream);
int cg_mlp_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                int step, void *stream);
int cg_tune_foreach(void (*fn)(void *ctx, const int *key, int n, const int *cfg, const int *splits,
                               const float *us),
                    void *ctx);
void cg_key_of(int elem, const cg_kparams *kp, int *k);

typedef struct {
    int ready;
    float mu[CM_F], sd[CM_F];
    float tmu, tsd;
    float params[CM_NP];
} cmodel_t;

static cmodel_t g_model;
static pthread_mutex_t m_lock = PTHREAD_MUTEX_INITIALIZER;

/* per-shape cache of the model's choice */
#define CACHE_N 4096
static struct { int used; int key[16]; int cfg, splits; } g_cache[CACHE_N];

/* ------------------------------------------------------------- features */

static int esz(int elem) { return elem == CG_ELEM_F32 ? 4 : elem == CG_ELEM_BF16 ? 2 : 1; }
static float lg(double x) { return (float)log2(x > 1.0 ? x : 1.0); }

/* f[CM_F] for problem kp (shape fields set) run with (cfg, splits); 0 on success. */
int cg_features(int elem, const cg_kparams *kp0, int cfg, int splits, float *f) {
    cg_tilecfg t;
    if (cg_get_tilecfg(elem, cfg, &t)) return -1;
    cg_kparams kp = *kp0;
    cg_fill_tiles(&kp, &t, splits > 0 ? splits : 1);
    memset(f, 0, sizeof(float) * CM_F);
    f[kp.op] = 1.0f;
    f[6 + elem] = 1.0f;
    f[9] = (float)kp.vec;
    f[10] = lg((double)kp.P);
    f[11] = lg(kp.C);
    f[12] = lg(kp.KN);
    f[13] = lg(kp.taps);
    f[14] = lg((double)kp.vs * kp.hs);
    f[15] = lg(kp.M);
    f[16] = lg(kp.N);
    f[17] = lg((double)kp.KT * t.BK);
    f[18] = lg(t.BM);
    f[19] = lg(t.BN);
    f[20] = lg((double)t.BK * esz(elem));
    f[21] = lg(kp.splits);
    long long wgs = (long long)kp.mtiles * kp.ntiles * kp.splits * kp.ncls_h * kp.ncls_w;
    f[22] = lg((double)wgs);
    int lds = t.lp ? (t.BM + t.BN) * 128 * t.lp : (t.BM + t.BN) * (t.BK * esz(elem) + 16) * 2;
    int per_cu = 163840 / lds;
    if (per_cu < 1) per_cu = 1;
    if (per_cu > 8) per_cu = 8;
    f[23] = (float)((double)kp.M * kp.N / ((double)kp.mtiles * t.BM * kp.ntiles * t.BN));
    f[24] = lg((double)((wgs + 256LL * per_cu - 1) / (256LL * per_cu)));
    f[25] = lg(2.0 * (double)kp.M * kp.N * (double)kp.KT * t.BK);
    f[26] = t.lp ? 1.0f : 0.0f;
    f[28] = lg(t.lp ? (double)t.WM * t.WN : 4.0); /* waves per workgroup (4 or 8) */
    f[29] = (float)t.lp;                           /* LDS stages of the pipelined kernels */
    f[30] = t.raster == 1 ? 1.0f : 0.0f;           /* raster order: M panels adjacent */
    f[31] = t.raster == 2 ? 1.0f : 0.0f;           /*               bands of 8 M panels */
    f[32] = t.halo ? 1.0f : 0.0f;                  /* the halo-tile kernel */
    if (splits == 0) { /* stream-K: data-parallel tiles + stream-K workgroups, no split partials */
        int dp;
        const int g = cg_sk_split(elem, cfg, &kp, &dp);
        f[22] = lg((double)g + dp);
        f[24] = 0.0f;
        f[27] = 1.0f;
    }
    return 0;
}

/* Rebuild the kernel parameters of a tuning-database key. */
static void kp_of_key(const int *k, cg_kparams *kp, int *elem) {
    memset(kp, 0, sizeof(*kp));
    *elem = k[0];
    kp->op = k[1]; kp->vec = k[2]; kp->B = k[3]; kp->H = k[4]; kp->W = k[5]; kp->C = k[6]; kp->KN = k[7];
    kp->KH = k[8]; kp->KW = k[9]; kp->vp = k[10]; kp->hp = k[11]; kp->vs = k[12]; kp->hs = k[13];
    kp->vd = k[14]; kp->hd = k[15];
    kp->Ho = convgemm_out_size(kp->H, kp->vp, kp->vd, kp->KH, kp->vs);
    kp->Wo = convgemm_out_size(kp->W, kp->hp, kp->hd, kp->KW, kp->hs);
    kp->taps = kp->KH * kp->KW;
    if (cg_dgrad_forward_form_ok(kp)) cg_forward_form(kp); /* as the call ran it */
    kp->P = (long long)kp->B * kp->Ho * kp->Wo;
}

/* ---------------------------------------------------------- device MLP */

typedef struct {
    int N;
    float *X, *T, *P, *G, *M, *V, *H1, *H2, *Y, *dY, *dZ2, *dZ1, *part;
} dbufs;

static void dfree(dbufs *d) {
    float *ps[] = {d->X, d->T, d->P, d->G, d->M, d->V, d->H1, d->H2, d->Y, d->dY, d->dZ2, d->dZ1, d->part};
    for (unsigned i = 0; i < sizeof(ps) / sizeof(ps[0]); i++)
        if (ps[i]) hipFree(ps[i]);
    memset(d, 0, sizeof(*d));
}

static int dalloc(dbufs *d, int N, int train) {
    memset(d, 0, sizeof(*d));
    d->N = N;
    int nb = (N + 255) / 256;
    hipError_t e = hipSuccess;
    e = e ? e : hipMalloc((void **)&d->X, sizeof(float) * (size_t)N * CM_F);
    e = e ? e : hipMalloc((void **)&d->P, sizeof(float) * CM_NP);
    e = e ? e : hipMalloc((void **)&d->H1, sizeof(float) * (size_t)N * CM_H);
    e = e ? e : hipMalloc((void **)&d->H2, sizeof(float) * (size_t)N * CM_H);
    e = e ? e : hipMalloc((void **)&d->Y, sizeof(float) * (size_t)N);
    if (train) {
        e = e ? e : hipMalloc((void **)&d->T, sizeof(float) * (size_t)N);
        e = e ? e : hipMalloc((void **)&d->G, sizeof(float) * CM_NP);
        e = e ? e : hipMalloc((void **)&d->M, sizeof(float) * CM_NP);
        e = e ? e : hipMalloc((void **)&d->V, sizeof(float) * CM_NP);
        e = e ? e : hipMalloc((void **)&d->dY, sizeof(float) * (size_t)N);
        e = e ? e : hipMalloc((void **)&d->dZ2, sizeof(float) * (size_t)N * CM_H);
        e = e ? e : hipMalloc((void **)&d->dZ1, sizeof(float) * (size_t)N * CM_H);
        e = e ? e : hipMalloc((void **)&d->part, sizeof(float) * (size_t)nb);
    }
    if (e != hipSuccess) {
        dfree(d);
        return -1;
    }
    return 0;
}

static int forward(dbufs *d, hipStream_t s) {
    int e = 0;
    e = e ? e : cg_mlp_dense_fwd(d->X, d->P + OFF_W1, d->P + OFF_B1, d->H1, d->N, CM_F, CM_H, 1, s);
    e = e ? e : cg_mlp_dense_fwd(d->H1, d->P + OFF_W2, d->P + OFF_B2, d->H2, d->N, CM_H, CM_H, 1, s);
    e = e ? e : cg_mlp_dense_fwd(d->H2, d->P + OFF_W3, d->P + OFF_B3, d->Y, d->N, CM_H, 1, 0, s);
    return e;
}

static int backward(dbufs *d, hipStream_t s) {
    int e = 0, N = d->N;
    e = e ? e : cg_mlp_mse_grad(d->Y, d->T, d->dY, d->part, N, s);
    e = e ? e : cg_mlp_dense_wgrad(d->H2, d->dY, d->G + OFF_W3, d->G + OFF_B3, N, CM_H, 1, s);
    e = e ? e : cg_mlp_dense_bwd_act(d->dY, d->P + OFF_W3, d->H2, d->dZ2, N, CM_H, 1, s);
    e = e ? e : cg_mlp_dense_wgrad(d->H1, d->dZ2, d->G + OFF_W2, d->G + OFF_B2, N, CM_H, CM_H, s);
    e = e ? e : cg_mlp_dense_bwd_act(d->dZ2, d->P + OFF_W2, d->H1, d->dZ1, N, CM_H, CM_H, s);
    e = e ? e : cg_mlp_dense_wgrad(d->X, d->dZ1, d->G + OFF_W1, d->G + OFF_B1, N, CM_F, CM_H, s);
    return e;
}

static float read_loss(dbufs *d, hipStream_t s) {
    int nb = (d->N + 255) / 256;
    float *h = malloc(sizeof(float) * (size_t)nb);
    if (!h) return NAN;
    hipMemcpyAsync(h, d->part, sizeof(float) * (size_t)nb, hipMemcpyDeviceToHost, s);
    hipStreamSynchronize(s);
    double t = 0;
    for (int i = 0; i < nb; i++) t += h[i];
    free(h);
    return (float)(t / d->N);
}

/* Test hook: loss and gradients of the MLP at `params` on (X, T), computed by
 * the device kernels.  X [N][CM_F], T [N], grads [CM_NP]. */
int cg_mlp_loss_grad(const float *params, const float *X, const float *T, int N, float *grads, float *loss) {
    dbufs d;
    if (N <= 0 || dalloc(&d, N, 1)) return -1;
    hipStream_t s = 0;
    hipMemcpy(d.P, params, sizeof(float) * CM_NP, hipMemcpyHostToDevice);
    hipMemcpy(d.X, X, sizeof(float) * (size_t)N * CM_F, hipMemcpyHostToDevice);
    hipMemcpy(d.T, T, sizeof(float) * (size_t)N, hipMemcpyHostToDevice);
    int e = forward(&d, s);
    e = e ? e : backward(&d, s);
    if (!e) {
        *loss = read_loss(&d, s);
        hipMemcpy(grads, d.G, sizeof(float) * CM_NP, hipMemcpyDeviceToHost);
    }
    dfree(&d);
    return e;
}

int cg_mlp_dims(int *F, int *H, int *NP) {
    *F = CM_F;
    *H = CM_H;
    *NP = CM_NP;
    return 0;
}

/* ---------------------------------------------------------- training set */

typedef struct {
    float *X, *T;
    int n, cap;
} tset;

static void collect(void *ctx, const int *key, int n, const int *cfg, const int *splits, const float *us) {
    tset *ts = ctx;
    cg_kparams kp;
    int elem;
    kp_of_key(key, &kp, &elem);
    for (int i = 0; i < n; i++) {
        if (!(us[i] > 0.0f) || us[i] >= 1e29f) continue;
        if (ts->n == ts->cap) {
            int nc = ts->cap ? ts->cap * 2 : 4096;
            float *nx = realloc(ts->X, sizeof(float) * (size_t)nc * CM_F);
            if (!nx) return;
            ts->X = nx;
            float *nt = realloc(ts->T, sizeof(float) * (size_t)nc);
            if (!nt) return;
            ts->T = nt;
            ts->cap = nc;
        }
        if (cg_features(elem, &kp, cfg[i], splits[i], ts->X + (size_t)ts->n * CM_F)) continue;
        ts->T[ts->n] = logf(us[i]);
        ts->n++;
    }
}

static uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double gauss(uint64_t *s) {
    double u1 = ((sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double u2 = ((sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static convgemm_train_progress_t g_progress;
static int g_progress_every;
static void *g_progress_ctx;

void convgemm_costmodel_set_progress(convgemm_train_progress_t fn, int every, void *ctx) {
    pthread_mutex_lock(&m_lock);
    g_progress = fn;
    g_progress_every = every;
    g_progress_ctx = ctx;
    pthread_mutex_unlock(&m_lock);
}

convgemm_status_t convgemm_costmodel_train(int epochs, float lr, unsigned seed, float *loss,
                                           convgemm_stream_t stream) {
    pthread_mutex_lock(&m_lock);
    const convgemm_train_progress_t progress = g_progress;
    const int every = g_progress_every;
    void *const pctx = g_progress_ctx;
    pthread_mutex_unlock(&m_lock);
    tset ts = {0};
    cg_tune_foreach(collect, &ts);
    if (ts.n < 8) {
        free(ts.X);
        free(ts.T);
        return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "cost model: fewer than 8 timings in the tuning database");
    }
    cmodel_t m;
    memset(&m, 0, sizeof(m));
    /* standardise features and target */
    for (int j = 0; j < CM_F; j++) {
        double s = 0, s2 = 0;
        for (int i = 0; i < ts.n; i++) {
            double v = ts.X[(size_t)i * CM_F + j];
            s += v;
            s2 += v * v;
        }
        double mu = s / ts.n, var = s2 / ts.n - mu * mu;
        m.mu[j] = (float)mu;
        m.sd[j] = (float)(var > 1e-12 ? sqrt(var) : 1.0);
    }
    double s = 0, s2 = 0;
    for (int i = 0; i < ts.n; i++) {
        s += ts.T[i];
        s2 += (double)ts.T[i] * ts.T[i];
    }
    m.tmu = (float)(s / ts.n);
    double tv = s2 / ts.n - (double)m.tmu * m.tmu;
    m.tsd = (float)(tv > 1e-12 ? sqrt(tv) : 1.0);
    for (int i = 0; i < ts.n; i++) {
        for (int j = 0; j < CM_F; j++) {
            float *v = &ts.X[(size_t)i * CM_F + j];
            *v = (*v - m.mu[j]) / m.sd[j];
        }
        ts.T[i] = (ts.T[i] - m.tmu) / m.tsd;
    }
    /* He initialisation, zero biases */
    uint64_t rs = 0x5EED0000ull ^ seed;
    for (int i = 0; i < CM_F * CM_H; i++) m.params[OFF_W1 + i] = (float)(gauss(&rs) * sqrt(2.0 / CM_F));
    for (int i = 0; i < CM_H * CM_H; i++) m.params[OFF_W2 + i] = (float)(gauss(&rs) * sqrt(2.0 / CM_H));
    for (int i = 0; i < CM_H; i++) m.params[OFF_W3 + i] = (float)(gauss(&rs) * sqrt(1.0 / CM_H));

    dbufs d;
    if (dalloc(&d, ts.n, 1)) {
        free(ts.X);
        free(ts.T);
        return cg_set_error(CONVGEMM_ERR_ALLOC, "cost model: device allocation failed");
    }
    hipStream_t st = (hipStream_t)stream;
    hipMemcpyAsync(d.X, ts.X, sizeof(float) * (size_t)ts.n * CM_F, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(d.T, ts.T, sizeof(float) * (size_t)ts.n, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(d.P, m.params, sizeof(float) * CM_NP, hipMemcpyHostToDevice, st);
    hipMemsetAsync(d.M, 0, sizeof(float) * CM_NP, st);
    hipMemsetAsync(d.V, 0, sizeof(float) * CM_NP, st);
    int e = 0;
    if (epochs < 1) epochs = 1;
    for (int ep = 1; ep <= epochs && !e; ep++) {
        e = forward(&d, st);
        e = e ? e : backward(&d, st);
        /* the loss partials of this epoch's parameters are in d.part after backward */
        if (!e && progress && every > 0 && ep % every == 0) progress(ep, read_loss(&d, st), pctx);
        e = e ? e : cg_mlp_adam(d.P, d.G, d.M, d.V, CM_NP, lr, 0.9f, 0.999f, 1e-8f, ep, st);
    }
    float final_loss = NAN;
    if (!e) {
        e = forward(&d, st);
        e = e ? e : cg_mlp_mse_grad(d.Y, d.T, d.dY, d.part, d.N, st);
        final_loss = read_loss(&d, st);
        hipMemcpyAsync(m.params, d.P, sizeof(float) * CM_NP, hipMemcpyDeviceToHost, st);
        hipStreamSynchronize(st);
    }
    dfree(&d);
    free(ts.X);
    free(ts.T);
    if (e) return cg_set_error(CONVGEMM_ERR_HIP, "cost model: kernel launch failed");
    m.ready = 1;
    pthread_mutex_lock(&m_lock);
    g_model = m;
    memset(g_cache, 0, sizeof(g_cache));
    pthread_mutex_unlock(&m_lock);
    if (loss) *loss = final_loss;
    return cg_set_error(CONVGEMM_SUCCESS, "");
}

/* ------------------------------------------------------------ inference */

/* Predicted microseconds of candidates (cfgs[i], splits[i]) of problem kp. */
static int predict(int elem, const cg_kparams *kp, const int *cfgs, const int *splits, int n, float *us) {
    cmodel_t m;
    pthread_mutex_lock(&m_lock);
    m = g_model;
    pthread_mutex_unlock(&m_lock);
    if (!m.ready || n <= 0) return -1;
    float *X = malloc(sizeof(float) * (size_t)n * CM_F);
    if (!X) return -1;
    for (int i = 0; i < n; i++) {
        float *f = X + (size_t)i * CM_F;
        if (cg_features(elem, kp, cfgs[i], splits[i], f)) memset(f, 0, sizeof(float) * CM_F);
        for (int j = 0; j < CM_F; j++) f[j] = (f[j] - m.mu[j]) / m.sd[j];
    }
    dbufs d;
    if (dalloc(&d, n, 0)) {
        free(X);
        return -1;
    }
    hipMemcpy(d.X, X, sizeof(float) * (size_t)n * CM_F, hipMemcpyHostToDevice);
    hipMemcpy(d.P, m.params, sizeof(float) * CM_NP, hipMemcpyHostToDevice);
    int e = forward(&d, 0);
    if (!e) e = hipMemcpy(us, d.Y, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess;
    dfree(&d);
    free(X);
    if (e) return -1;
    for (int i = 0; i < n; i++) us[i] = expf(us[i] * m.tsd + m.tmu);
    return 0;
}

/* Predicted times of n candidates of one problem; 0, or -1 without a model
 * (runs the model's device kernels: not for a capturing stream). */
int cg_costmodel_times(int elem, const cg_kparams *kp, const int *cfgs, const int *splits, int n, float *us) {
    pthread_mutex_lock(&m_lock);
    const int ready = g_model.ready;
    pthread_mutex_unlock(&m_lock);
    if (!ready || n <= 0) return -1;
    return predict(elem, kp, cfgs, splits, n, us) ? -1 : 0;
}

int cg_costmodel_choose(int elem, const cg_kparams *kp, int *cfg, int *splits, int allow_device) {
    int key[16];
    cg_key_of(elem, kp, key);
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 16; i++) h = (h ^ (uint32_t)key[i]) * 1099511628211ull;
    int slot = (int)(h % CACHE_N);
    pthread_mutex_lock(&m_lock);
    int ready = g_model.ready;
    if (ready && g_cache[slot].used && !memcmp(g_cache[slot].key, key, sizeof(key))) {
        *cfg = g_cache[slot].cfg;
        *splits = g_cache[slot].splits;
        pthread_mutex_unlock(&m_lock);
        return 0;
    }
    pthread_mutex_unlock(&m_lock);
    if (!ready || !allow_device) return -1;
    int cfgs[CONVGEMM_MAX_CANDIDATES], spl[CONVGEMM_MAX_CANDIDATES];
    float us[CONVGEMM_MAX_CANDIDATES];
    int n = cg_candidates(elem, kp, cfgs, spl, CONVGEMM_MAX_CANDIDATES);
    if (n <= 0 || predict(elem, kp, cfgs, spl, n, us)) return -1;
    int bi = 0;
    for (int i = 1; i < n; i++)
        if (us[i] < us[bi]) bi = i;
    *cfg = cfgs[bi];
    *splits = spl[bi];
    pthread_mutex_lock(&m_lock);
    g_cache[slot].used = 1;
    memcpy(g_cache[slot].key, key, sizeof(key));
    g_cache[slot].cfg = *cfg;
    g_cache[slot].splits = *splits;
    pthread_mutex_unlock(&m_lock);
    return 0;
}

convgemm_status_t convgemm_costmodel_predict(const convgemm_desc_t *desc, convgemm_dtype_t dtype, int *cfgs,
                                             int *splits, float *us, int *n_out) {
    cg_kparams kp;
    int elem, chunk;
    convgemm_status_t st = cg_prepare(desc, dtype, (void *)256, (void *)256, NULL, &kp, &elem, &chunk);
    if (st != CONVGEMM_SUCCESS) return st;
    if (!g_model.ready) return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "cost model not trained or loaded");
    int c[CONVGEMM_MAX_CANDIDATES], s[CONVGEMM_MAX_CANDIDATES];
    float u[CONVGEMM_MAX_CANDIDATES];
    int n = cg_candidates(elem, &kp, c, s, CONVGEMM_MAX_CANDIDATES);
    if (n > *n_out) n = *n_out;
    if (predict(elem, &kp, c, s, n, u)) return cg_set_error(CONVGEMM_ERR_HIP, "cost model inference failed");
    for (int i = 0; i < n; i++) {
        cfgs[i] = c[i];
        splits[i] = s[i];
        us[i] = u[i];
    }
    *n_out = n;
    return cg_set_error(CONVGEMM_SUCCESS, "");
}

/* -------------------------------------------------------------- files */

static const char MAGIC[8] = {'C', 'G', 'M', 'L', 'P', '0', '0', '4'};

convgemm_status_t convgemm_costmodel_save(const char *path) {
    pthread_mutex_lock(&m_lock);
    cmodel_t m = g_model;
    pthread_mutex_unlock(&m_lock);
    if (!m.ready) return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "cost model not trained or loaded");
    FILE *f = fopen(path, "wb");
    if (!f) return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "cannot open cost model file for writing");
    int dims[3] = {CM_F, CM_H, CM_NP};
    int ok = fwrite(MAGIC, 1, 8, f) == 8 && fwrite(dims, sizeof(int), 3, f) == 3 &&
             fwrite(m.mu, sizeof(float), CM_F, f) == CM_F && fwrite(m.sd, sizeof(float), CM_F, f) == CM_F &&
             fwrite(&m.tmu, sizeof(float), 1, f) == 1 && fwrite(&m.tsd, sizeof(float), 1, f) == 1 &&
             fwrite(m.params, sizeof(float), CM_NP, f) == CM_NP;
    fclose(f);
    return ok ? cg_set_error(CONVGEMM_SUCCESS, "") : cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "write error");
}

convgemm_status_t convgemm_costmodel_load(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "cannot open cost model file");
    cmodel_t m;
    memset(&m, 0, sizeof(m));
    char mg[8];
    int dims[3];
    int ok = fread(mg, 1, 8, f) == 8 && !memcmp(mg, MAGIC, 8) && fread(dims, sizeof(int), 3, f) == 3;
    /* this build's feature count, or an earlier one: its first-layer rows for the later
     * features are zero (inputs the network ignores) */
    const int F = ok ? dims[0] : 0;
    ok = ok && F >= CM_F_V1 && F <= CM_F && dims[1] == CM_H &&
         dims[2] == F * CM_H + CM_H + CM_H * CM_H + CM_H + CM_H + 1;
    for (int j = 0; j < CM_F; j++) {
        m.mu[j] = 0.0f;
        m.sd[j] = 1.0f;
    }
    ok = ok && fread(m.mu, sizeof(float), F, f) == (size_t)F && fread(m.sd, sizeof(float), F, f) == (size_t)F &&
         fread(&m.tmu, sizeof(float), 1, f) == 1 && fread(&m.tsd, sizeof(float), 1, f) == 1 &&
         fread(m.params, sizeof(float), (size_t)F * CM_H, f) == (size_t)F * CM_H &&
         fread(m.params + OFF_B1, sizeof(float), CM_NP - OFF_B1, f) == (size_t)(CM_NP - OFF_B1);
    fclose(f);
    if (!ok) return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "not a convgemm cost model file");
    m.ready = 1;
    pthread_mutex_lock(&m_lock);
    g_model = m;
    memset(g_cache, 0, sizeof(g_cache));
    pthread_mutex_unlock(&m_lock);
    return cg_set_error(CONVGEMM_SUCCESS, "");
}
