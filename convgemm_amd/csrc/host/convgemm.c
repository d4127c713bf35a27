/*
 * convgemm.c -- C host layer of the MI355X implicit-GEMM convolution library.
 *
 * Exports the reference's five entry points (src/convGemm.h:20-64) on device
 * pointers plus the extended API of include/convgemm.h.  All device work is
 * reached through the extern "C" launchers of cg_internal.h.
 *
 * Per call: validate -> reference quick-return rules -> plan (tile
 * configuration + split-K) -> optional weight repack -> implicit GEMM ->
 * optional split-K reduction.  Drop-in calls then wait for the stream.
 */
#include <pthread.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "convgemm.h"
#include "../cg_internal.h"
#include "cg_host.h"

#define CG_VERSION 10000 /* 1.0.0 */

/* ------------------------------------------------------------------ state */

static _Thread_local convgemm_status_t t_last_status = CONVGEMM_SUCCESS;
static _Thread_local char t_last_msg[256] = "";

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static hipStream_t g_stream = NULL;

/* internal workspace used when the caller passes none */
static void *g_ws = NULL;
static size_t g_ws_bytes = 0;

/* buffers handed out by alloc_pack_buffs, with their sizes */
#define CG_MAX_PACKS 64
static struct { void *ptr; size_t bytes; } g_packs[CG_MAX_PACKS];

/* forced plan (tuning hook) */
static int g_force_cfg = -1, g_force_splits = 0;

/* optional planner hook installed by the autotuner / cost model */
static cg_plan_hook_t g_plan_hook = NULL;
static void *g_plan_hook_ctx = NULL;

static convgemm_status_t set_error(convgemm_status_t st, const char *fmt, ...) {
    t_last_status = st;
    if (fmt) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(t_last_msg, sizeof(t_last_msg), fmt, ap);
        va_end(ap);
    } else {
        t_last_msg[0] = 0;
    }
    return st;
}

convgemm_status_t cg_set_error(convgemm_status_t st, const char *msg) { return set_error(st, "%s", msg); }

static convgemm_status_t ok(void) {
    t_last_status = CONVGEMM_SUCCESS;
    t_last_msg[0] = 0;
    return CONVGEMM_SUCCESS;
}

convgemm_status_t convgemm_get_last_error(void) { return t_last_status; }
const char *convgemm_last_error_message(void) { return t_last_msg; }

const char *convgemm_status_string(convgemm_status_t s) {
    switch (s) {
    case CONVGEMM_SUCCESS: return "success";
    case CONVGEMM_ERR_INVALID_VALUE: return "invalid value";
    case CONVGEMM_ERR_ALLOC: return "device allocation failed";
    case CONVGEMM_ERR_HIP: return "HIP runtime error";
    case CONVGEMM_ERR_WORKSPACE: return "workspace too small";
    case CONVGEMM_ERR_UNSUPPORTED: return "unsupported problem";
    }
    return "unknown status";
}

int convgemm_version(void) { return CG_VERSION; }

convgemm_status_t convgemm_set_stream(convgemm_stream_t stream) {
    pthread_mutex_lock(&g_lock);
    g_stream = (hipStream_t)stream;
    pthread_mutex_unlock(&g_lock);
    return ok();
}

void cg_force_config(int cfg, int splits) {
    g_force_cfg = cfg;
    g_force_splits = splits;
}

void cg_set_plan_hook(cg_plan_hook_t hook, void *ctx) {
    g_plan_hook = hook;
    g_plan_hook_ctx = ctx;
}

/* plan of the last call on this thread (introspection for tuning and benchmarks) */
static _Thread_local int t_plan_cfg = -1, t_plan_splits = 0, t_plan_kernels = 0;

void cg_last_plan(int *cfg, int *splits, int *nkernels) {
    if (cfg) *cfg = t_plan_cfg;
    if (splits) *splits = t_plan_splits;
    if (nkernels) *nkernels = t_plan_kernels;
}

/* Test-only: make the next `n` device allocations of the library fail. */
static int g_fail_allocs = 0;
void cg_debug_fail_next_allocs(int n) { g_fail_allocs = n; }

static hipError_t lib_malloc(void **p, size_t bytes) {
    if (g_fail_allocs > 0) {
        g_fail_allocs--;
        *p = NULL;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

/* ------------------------------------------------------------- shapes */

int convgemm_out_size(int in, int pad, int dil, int k, int stride) {
    if (stride <= 0) return 0;
    return (in + 2 * pad - dil * (k - 1) - 1) / stride + 1;
}

static convgemm_status_t validate(const convgemm_desc_t *d, int *ho, int *wo) {
    if (!d) return set_error(CONVGEMM_ERR_INVALID_VALUE, "descriptor is NULL");
    if (d->op < 0 || d->op > 2) return set_error(CONVGEMM_ERR_INVALID_VALUE, "unknown op %d", d->op);
    if (d->layout < 0 || d->layout > 1)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "unknown layout %d", d->layout);
    if (d->b < 0 || d->h < 0 || d->w < 0 || d->c < 0 || d->kn < 0)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "negative dimension (b=%d h=%d w=%d c=%d kn=%d)",
                         d->b, d->h, d->w, d->c, d->kn);
    if (d->kh < 1 || d->kw < 1)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "kernel size must be >= 1 (kh=%d kw=%d)", d->kh, d->kw);
    if (d->vstride < 1 || d->hstride < 1)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "stride must be >= 1 (%d, %d)", d->vstride, d->hstride);
    if (d->vdil < 1 || d->hdil < 1)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "dilation must be >= 1 (%d, %d)", d->vdil, d->hdil);
    if (d->vpad < 0 || d->hpad < 0)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "padding must be >= 0 (%d, %d)", d->vpad, d->hpad);
    *ho = convgemm_out_size(d->h, d->vpad, d->vdil, d->kh, d->vstride);
    *wo = convgemm_out_size(d->w, d->hpad, d->hdil, d->kw, d->hstride);
    if (*ho <= 0 || *wo <= 0)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "empty output (ho=%d wo=%d): kernel larger than padded input",
                         *ho, *wo);
    /* 32-bit index ranges used by the kernels */
    long long P = (long long)d->b * *ho * *wo;
    long long X = (long long)d->b * d->h * d->w * d->c;
    long long K = (long long)d->c * d->kh * d->kw;
    if (P >= (1ll << 31) || X >= (1ll << 40) || K * d->kn >= (1ll << 40) || K >= (1ll << 31) ||
        (long long)d->h * d->w >= (1ll << 31))
        return set_error(CONVGEMM_ERR_UNSUPPORTED, "problem too large for 32-bit pixel indexing");
    if (d->op == CONVGEMM_DGRAD && (d->kh > 64 || d->kw > 64))
        return set_error(CONVGEMM_ERR_UNSUPPORTED, "backward data supports kernels up to 64x64");
    return CONVGEMM_SUCCESS;
}

/* --------------------------------------------------------- workspace */

static size_t pack_bytes_of(const void *p) {
    if (!p) return 0;
    size_t r = 0;
    pthread_mutex_lock(&g_lock);
    for (int i = 0; i < CG_MAX_PACKS; i++)
        if (g_packs[i].ptr == p) r = g_packs[i].bytes;
    pthread_mutex_unlock(&g_lock);
    /* unknown caller buffer: the reference contract says it is at least as
     * large as what alloc_pack_buffs hands out */
    return r ? r : CG_PACK_BYTES;
}

static convgemm_status_t internal_ws(size_t need, void **out) {
    pthread_mutex_lock(&g_lock);
    if (need > g_ws_bytes) {
        /* the old buffer may still be read by queued work on any stream */
        if (g_ws) {
            hipDeviceSynchronize();
            hipFree(g_ws);
        }
        g_ws = NULL;
        g_ws_bytes = 0;
        size_t sz = need + (need >> 2);
        if (lib_malloc(&g_ws, sz) != hipSuccess) {
            g_ws = NULL;
            pthread_mutex_unlock(&g_lock);
            return set_error(CONVGEMM_ERR_ALLOC, "cannot allocate %zu bytes of internal workspace", sz);
        }
        g_ws_bytes = sz;
    }
    *out = g_ws;
    pthread_mutex_unlock(&g_lock);
    return CONVGEMM_SUCCESS;
}

int alloc_pack_buffs(float **Ac_pack, float **Bc_pack) {
    if (!Ac_pack || !Bc_pack) {
        set_error(CONVGEMM_ERR_INVALID_VALUE, "NULL output pointer");
        return 1;
    }
    *Ac_pack = NULL;
    *Bc_pack = NULL;
    void *a = NULL, *b = NULL;
    if (lib_malloc(&a, CG_PACK_BYTES) != hipSuccess) a = NULL;
    if (a && lib_malloc(&b, CG_PACK_BYTES) != hipSuccess) b = NULL;
    if (!a || !b) {
        if (a) hipFree(a);
        set_error(CONVGEMM_ERR_ALLOC, "cannot allocate pack buffers (2 x %d bytes)", CG_PACK_BYTES);
        return 1;
    }
    pthread_mutex_lock(&g_lock);
    int placed = 0;
    for (int i = 0; i < CG_MAX_PACKS && placed < 2; i++)
        if (!g_packs[i].ptr) {
            g_packs[i].ptr = placed == 0 ? a : b;
            g_packs[i].bytes = CG_PACK_BYTES;
            placed++;
        }
    pthread_mutex_unlock(&g_lock);
    *Ac_pack = (float *)a;
    *Bc_pack = (float *)b;
    ok();
    return 0;
}

void convgemm_free_pack_buffs(float *Ac_pack, float *Bc_pack) {
    void *ps[2] = {Ac_pack, Bc_pack};
    pthread_mutex_lock(&g_lock);
    for (int k = 0; k < 2; k++) {
        if (!ps[k]) continue;
        for (int i = 0; i < CG_MAX_PACKS; i++)
            if (g_packs[i].ptr == ps[k]) g_packs[i].ptr = NULL;
    }
    pthread_mutex_unlock(&g_lock);
    hipDeviceSynchronize();
    if (Ac_pack) hipFree(Ac_pack);
    if (Bc_pack) hipFree(Bc_pack);
}

/* ------------------------------------------------------------- planning */

static int opkind(const convgemm_desc_t *d) {
    if (d->op == CONVGEMM_FWD) return d->layout == CONVGEMM_NHWC ? CG_FWD_NHWC : CG_FWD_NCHW;
    if (d->op == CONVGEMM_WGRAD) return d->layout == CONVGEMM_NHWC ? CG_WGRAD_NHWC : CG_WGRAD_NCHW;
    return d->layout == CONVGEMM_NHWC ? CG_DGRAD_NHWC : CG_DGRAD_NCHW;
}

static int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

/* Fill the tile-space fields of kp for configuration cfg. */
void cg_fill_tiles(cg_kparams *kp, const cg_tilecfg *t, int splits) {
    int op = kp->op;
    kp->ncls_h = 1;
    kp->ncls_w = 1;
    if (op == CG_FWD_NHWC || op == CG_FWD_NCHW) {
        kp->M = (int)kp->P;
        kp->N = kp->KN;
        kp->kchunks = cdiv(kp->C, t->BK);
        kp->KT = kp->taps * kp->kchunks;
    } else if (op == CG_DGRAD_NHWC || op == CG_DGRAD_NCHW) {
        kp->ncls_h = kp->vs;
        kp->ncls_w = kp->hs;
        int hc = cdiv(kp->H, kp->vs), wc = cdiv(kp->W, kp->hs); /* class (0,0) is the largest */
        kp->M = kp->B * hc * wc;
        kp->N = kp->C;
        kp->kchunks = cdiv(kp->KN, t->BK);
        kp->KT = kp->taps * kp->kchunks;
    } else {
        kp->M = kp->taps * kp->C;
        kp->N = kp->KN;
        kp->kchunks = cdiv(kp->C, t->BM); /* M chunks per tap */
        kp->KT = cdiv(kp->P, t->BK);
    }
    if (op == CG_WGRAD_NHWC || op == CG_WGRAD_NCHW)
        kp->mtiles = kp->taps * kp->kchunks;
    else
        kp->mtiles = cdiv(kp->M, t->BM);
    kp->ntiles = cdiv(kp->N, t->BN);
    if (splits < 1) splits = 1;
    if (splits > kp->KT) splits = kp->KT > 0 ? kp->KT : 1;
    kp->splits = splits;
}

size_t cg_partial_bytes(const cg_kparams *kp) {
    if (kp->splits <= 1) return 0;
    return (size_t)kp->splits * kp->ncls_h * kp->ncls_w * (size_t)kp->M * kp->N * sizeof(float);
}

/* Analytic default plan: fill the 256 CUs, prefer large tiles, split K when
 * the tile grid is too small to occupy the chip. */
static void default_plan(const cg_kparams *base, int *cfg_out, int *splits_out) {
    const int CUS = 256;
    int op = base->op;
    int is_wg = (op == CG_WGRAD_NHWC || op == CG_WGRAD_NCHW);
    double best = 1e300;
    int best_cfg = 3, best_splits = 1;
    int n = cg_num_tilecfgs(CG_ELEM_F32);
    for (int cfg = 0; cfg < n; cfg++) {
        cg_tilecfg t;
        if (cg_get_tilecfg(CG_ELEM_F32, cfg, &t) || !cg_tilecfg_supports(CG_ELEM_F32, cfg, op)) continue;
        if (is_wg != (t.BK == 32)) continue;
        cg_kparams kp = *base;
        cg_fill_tiles(&kp, &t, 1);
        long long tiles = (long long)kp.mtiles * kp.ntiles * kp.ncls_h * kp.ncls_w;
        for (int s = 1; s <= 32; s *= 2) {
            if (s > 1 && kp.KT / s < 4) break;
            /* time ~ waves of workgroups x per-tile work, with a padding cost
             * for tile edges and a fixed cost per split-K reduction */
            long long wgs = tiles * s;
            int per_cu = 2; /* resident workgroups per CU */
            long long waves = (wgs + CUS * per_cu - 1) / (CUS * per_cu);
            double kwork = (double)kp.KT / s * t.BK;
            double tile_eff = (double)t.BM * t.BN / (t.BM + t.BN); /* reuse per loaded byte */
            double cost = waves * (double)t.BM * t.BN * kwork * per_cu / 1.0;
            cost *= (1.0 + 32.0 / tile_eff);
            if (s > 1) cost += (double)kp.M * kp.N * s * 0.5 * (kp.ncls_h * kp.ncls_w);
            if (cost < best) {
                best = cost;
                best_cfg = cfg;
                best_splits = s;
            }
        }
    }
    *cfg_out = best_cfg;
    *splits_out = best_splits;
}

/* ------------------------------------------------------------- execution */

typedef struct {
    void *ws_parts;
    size_t parts_bytes;
    void *ws_w;
    size_t w_bytes;
} cg_wsets;

static convgemm_status_t hip_fail(hipError_t e, const char *what) {
    return set_error(CONVGEMM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

static convgemm_status_t run(const convgemm_desc_t *d, float alpha, float beta, const void *in,
                             const void *x, void *out, const convgemm_epilogue_t *epi,
                             convgemm_dtype_t dtype, const float *scales, cg_wsets *ws,
                             hipStream_t stream) {
    int ho, wo;
    convgemm_status_t st = validate(d, &ho, &wo);
    if (st != CONVGEMM_SUCCESS) return st;
    if (dtype != CONVGEMM_F32) return cg_run_lowp(d, alpha, beta, in, x, out, epi, dtype, scales, stream);

    long long P = (long long)d->b * ho * wo;
    long long K = (long long)d->c * d->kh * d->kw;
    long long gm, gn, gk; /* the reference's GEMM orientation (src/convGemm.c) */
    long long out_elems;
    if (d->op == CONVGEMM_FWD) {
        gm = d->kn; gn = P; gk = K;
        out_elems = P * d->kn;
    } else if (d->op == CONVGEMM_WGRAD) {
        gm = d->kn; gn = K; gk = P;
        out_elems = K * d->kn;
    } else {
        gm = K; gn = P; gk = d->kn;
        out_elems = (long long)d->b * d->h * d->w * d->c;
    }
    /* quick return (src/gemm_blis_B3A2C0_orig.c:45-47) */
    if (gm == 0 || gn == 0) return ok();
    if ((alpha == 0.0f || gk == 0) && beta == 1.0f) return ok();
    if (!out) return set_error(CONVGEMM_ERR_INVALID_VALUE, "output pointer is NULL");
    if (alpha == 0.0f) { /* src/quick_gemm.h: C = beta*C, no epilogue */
        int e = cg_launch_scale((float *)out, out_elems, beta, stream);
        return e ? hip_fail((hipError_t)e, "scale kernel") : ok();
    }
    if (gk == 0) return ok(); /* the reference's K loop runs zero times: C untouched */
    if (!in || !x) return set_error(CONVGEMM_ERR_INVALID_VALUE, "input pointer is NULL");
    if (epi && epi->mean && (!epi->inv_std || !epi->gamma || !epi->beta))
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "batch norm needs inv_std, gamma and beta");

    cg_kparams kp;
    memset(&kp, 0, sizeof(kp));
    kp.op = opkind(d);
    kp.B = d->b; kp.H = d->h; kp.W = d->w; kp.C = d->c; kp.KN = d->kn; kp.KH = d->kh; kp.KW = d->kw;
    kp.vp = d->vpad; kp.hp = d->hpad; kp.vs = d->vstride; kp.hs = d->hstride; kp.vd = d->vdil; kp.hd = d->hdil;
    kp.Ho = ho; kp.Wo = wo; kp.taps = d->kh * d->kw;
    kp.P = P;
    kp.out = (float *)out;
    kp.alpha = alpha;
    kp.beta = beta;
    kp.scale = 1.0f;
    if (d->op == CONVGEMM_FWD && epi) {
        kp.bias = epi->bias; kp.mean = epi->mean; kp.inv_std = epi->inv_std;
        kp.gamma = epi->gamma; kp.bnb = epi->beta; kp.relu = epi->relu != 0;
    }
    /* operands */
    const float *wsrc = NULL;
    if (d->op == CONVGEMM_FWD) { kp.A = x; wsrc = (const float *)in; }
    else if (d->op == CONVGEMM_DGRAD) { kp.A = x; wsrc = (const float *)in; }
    else { kp.A = x; kp.Bm = in; }

    int need_wt = (d->op != CONVGEMM_WGRAD) && d->layout == CONVGEMM_NCHW;
    size_t wt_bytes = need_wt ? (size_t)K * d->kn * sizeof(float) : 0;
    int vec_ok_a = ((uintptr_t)kp.A % 16) == 0;
    int vec_ok_b = ((uintptr_t)(need_wt ? ws->ws_w : (const void *)wsrc) % 16) == 0 || d->op == CONVGEMM_WGRAD;
    if (kp.op == CG_FWD_NHWC) kp.vecA = vec_ok_a && d->c % 4 == 0;
    if (kp.op == CG_DGRAD_NHWC) kp.vecA = vec_ok_a && d->kn % 4 == 0;
    if (kp.op == CG_DGRAD_NHWC || kp.op == CG_DGRAD_NCHW) kp.vecB = vec_ok_b && d->kn % 4 == 0;

    /* plan */
    int cfg = -1, splits = 1;
    if (g_force_cfg >= 0) {
        cfg = g_force_cfg;
        splits = g_force_splits > 0 ? g_force_splits : 1;
    } else if (g_plan_hook) {
        g_plan_hook(g_plan_hook_ctx, &kp, &cfg, &splits);
    }
    if (cfg < 0 || !cg_tilecfg_supports(CG_ELEM_F32, cfg, kp.op)) default_plan(&kp, &cfg, &splits);
    cg_tilecfg t;
    if (cg_get_tilecfg(CG_ELEM_F32, cfg, &t))
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "unknown tile configuration %d", cfg);
    cg_fill_tiles(&kp, &t, splits);

    /* workspace */
    if (need_wt) {
        if (ws->ws_w && ws->w_bytes < wt_bytes)
            return set_error(CONVGEMM_ERR_WORKSPACE, "weight workspace %zu < %zu bytes", ws->w_bytes, wt_bytes);
    }
    size_t pbytes = cg_partial_bytes(&kp);
    while (pbytes && ws->ws_parts && pbytes > ws->parts_bytes && kp.splits > 1) {
        cg_fill_tiles(&kp, &t, kp.splits / 2);
        pbytes = cg_partial_bytes(&kp);
    }
    if ((need_wt && !ws->ws_w) || (pbytes && !ws->ws_parts)) {
        size_t wt_al = (wt_bytes + 255) & ~(size_t)255;
        void *base = NULL;
        st = internal_ws(wt_al + pbytes, &base);
        if (st != CONVGEMM_SUCCESS) return st;
        if (need_wt && !ws->ws_w) { ws->ws_w = base; ws->w_bytes = wt_bytes; }
        if (pbytes && !ws->ws_parts) { ws->ws_parts = (char *)base + wt_al; ws->parts_bytes = pbytes; }
    }
    if (need_wt) {
        /* NCHW weights [kn][K] -> [K][kn], the NHWC weight order the kernels read */
        int e = cg_launch_transpose_f32(wsrc, (float *)ws->ws_w, d->kn, (int)K, stream);
        if (e) return hip_fail((hipError_t)e, "weight transpose");
        kp.Bm = ws->ws_w;
        kp.vecB = kp.vecB && (((uintptr_t)ws->ws_w % 16) == 0);
    } else if (d->op != CONVGEMM_WGRAD) {
        kp.Bm = wsrc;
    }
    kp.ws = pbytes ? (float *)ws->ws_parts : NULL;

    int e = cg_launch_igemm(CG_ELEM_F32, cfg, &kp, stream);
    if (e) return hip_fail((hipError_t)e, "implicit GEMM launch");
    if (kp.splits > 1) {
        e = cg_launch_splitk_reduce(&kp, stream);
        if (e) return hip_fail((hipError_t)e, "split-K reduction");
    }
    t_plan_cfg = cfg;
    t_plan_splits = kp.splits;
    t_plan_kernels = 1 + (kp.splits > 1) + need_wt;
    return ok();
}

convgemm_status_t convgemm_conv_ex(const convgemm_desc_t *desc, float alpha, float beta, const void *in,
                                   const void *x, void *out, const convgemm_epilogue_t *epi,
                                   convgemm_dtype_t dtype, const float *scales, void *workspace, size_t ws_bytes,
                                   convgemm_stream_t stream) {
    cg_wsets ws = {0};
    if (workspace && ws_bytes) {
        /* weights first (256-byte aligned), partials after */
        int ho, wo;
        convgemm_status_t st = validate(desc, &ho, &wo);
        if (st != CONVGEMM_SUCCESS) return st;
        size_t wt = 0;
        if (desc->op != CONVGEMM_WGRAD && desc->layout == CONVGEMM_NCHW)
            wt = (size_t)desc->c * desc->kh * desc->kw * desc->kn * sizeof(float);
        size_t wt_al = (wt + 255) & ~(size_t)255;
        if (wt_al > ws_bytes)
            return set_error(CONVGEMM_ERR_WORKSPACE, "workspace %zu < %zu bytes needed for weights", ws_bytes, wt_al);
        ws.ws_w = wt ? workspace : NULL;
        ws.w_bytes = wt;
        ws.ws_parts = (char *)workspace + wt_al;
        ws.parts_bytes = ws_bytes - wt_al;
        if (!wt) ws.ws_w = workspace; /* unused */
        if (ws.parts_bytes == 0) ws.ws_parts = NULL;
    }
    return run(desc, alpha, beta, in, x, out, epi, dtype, scales, &ws, (hipStream_t)stream);
}

size_t convgemm_workspace_size(const convgemm_desc_t *d, convgemm_dtype_t dtype) {
    int ho, wo;
    if (validate(d, &ho, &wo) != CONVGEMM_SUCCESS) return 0;
    size_t wt = 0;
    if (d->op != CONVGEMM_WGRAD && d->layout == CONVGEMM_NCHW)
        wt = (size_t)d->c * d->kh * d->kw * d->kn * sizeof(float);
    size_t wt_al = (wt + 255) & ~(size_t)255;
    /* the largest partial buffer the default planner can choose */
    cg_kparams kp;
    memset(&kp, 0, sizeof(kp));
    kp.op = opkind(d);
    kp.B = d->b; kp.H = d->h; kp.W = d->w; kp.C = d->c; kp.KN = d->kn; kp.KH = d->kh; kp.KW = d->kw;
    kp.vp = d->vpad; kp.hp = d->hpad; kp.vs = d->vstride; kp.hs = d->hstride; kp.vd = d->vdil; kp.hd = d->hdil;
    kp.Ho = ho; kp.Wo = wo; kp.taps = d->kh * d->kw;
    kp.P = (long long)d->b * ho * wo;
    int cfg, splits;
    default_plan(&kp, &cfg, &splits);
    cg_tilecfg t;
    cg_get_tilecfg(CG_ELEM_F32, cfg, &t);
    cg_fill_tiles(&kp, &t, splits);
    (void)dtype;
    return wt_al + cg_partial_bytes(&kp) + cg_lowp_extra_bytes(d, dtype);
}

/* ------------------------------------------------------------- drop-in */

static void dropin(const convgemm_desc_t *d, float beta, const float *in, const float *x, float *out,
                   const convgemm_epilogue_t *epi, float *ac_pack, float *bc_pack) {
    cg_wsets ws = {0};
    /* Ac_pack holds split-K partials, Bc_pack the repacked weights */
    if (ac_pack) { ws.ws_parts = ac_pack; ws.parts_bytes = pack_bytes_of(ac_pack); }
    if (bc_pack) { ws.ws_w = bc_pack; ws.w_bytes = pack_bytes_of(bc_pack); }
    hipStream_t s = g_stream;
    convgemm_status_t st = run(d, 1.0f, beta, in, x, out, epi, CONVGEMM_F32, NULL, &ws, s);
    if (st != CONVGEMM_SUCCESS) return;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) hip_fail(e, "stream synchronize");
}

void sconvGemmNHWC(char trans, int b, int h, int w, int c, int kn, int kh, int kw, int vpadding, int hpadding,
                   int vstride, int hstride, int vdilation, int hdilation, const float *in, const float *x,
                   float *out, const float *bias_vector, const float *bn_running_mean, const float *bn_inv_std,
                   const float *bn_gamma, const float *bn_beta, bool relu, float *ac_pack, float *bc_pack) {
    convgemm_desc_t d = {trans == 'N' ? CONVGEMM_FWD : CONVGEMM_WGRAD, CONVGEMM_NHWC, b, h, w, c, kn, kh, kw,
                         vpadding, hpadding, vstride, hstride, vdilation, hdilation};
    convgemm_epilogue_t e = {bias_vector, bn_running_mean, bn_inv_std, bn_gamma, bn_beta, relu};
    dropin(&d, 0.0f, in, x, out, trans == 'N' ? &e : NULL, ac_pack, bc_pack);
}

void sconvGemmNHWC_back(int b, int h, int w, int c, int kn, int kh, int kw, int vstride, int hstride,
                        int vpadding, int hpadding, int vdilation, int hdilation, const float *weights,
                        const float *dy, float *dx, float *ac_pack, float *bc_pack) {
    convgemm_desc_t d = {CONVGEMM_DGRAD, CONVGEMM_NHWC, b, h, w, c, kn, kh, kw,
                         vpadding, hpadding, vstride, hstride, vdilation, hdilation};
    dropin(&d, 1.0f, weights, dy, dx, NULL, ac_pack, bc_pack);
}

void sconvGemmNCHW(char trans, int b, int c, int h, int w, int kn, int kh, int kw, int vpadding, int hpadding,
                   int vstride, int hstride, int vdilation, int hdilation, const float *in, const float *x,
                   float *out, const float *bias_vector, const float *bn_running_mean, const float *bn_inv_std,
                   const float *bn_gamma, const float *bn_beta, bool relu, float *ac_pack, float *bc_pack) {
    convgemm_desc_t d = {trans == 'N' ? CONVGEMM_FWD : CONVGEMM_WGRAD, CONVGEMM_NCHW, b, h, w, c, kn, kh, kw,
                         vpadding, hpadding, vstride, hstride, vdilation, hdilation};
    convgemm_epilogue_t e = {bias_vector, bn_running_mean, bn_inv_std, bn_gamma, bn_beta, relu};
    dropin(&d, 0.0f, in, x, out, trans == 'N' ? &e : NULL, ac_pack, bc_pack);
}

void sconvGemmNCHW_back(int b, int c, int h, int w, int kn, int kh, int kw, int vstride, int hstride,
                        int vpadding, int hpadding, int vdilation, int hdilation, const float *weights,
                        const float *dy, float *dx, float *ac_pack, float *bc_pack) {
    convgemm_desc_t d = {CONVGEMM_DGRAD, CONVGEMM_NCHW, b, h, w, c, kn, kh, kw,
                         vpadding, hpadding, vstride, hstride, vdilation, hdilation};
    dropin(&d, 1.0f, weights, dy, dx, NULL, ac_pack, bc_pack);
}
