/*
 * convgemm.c -- C host layer of the MI355X implicit-GEMM convolution library.
 *
 * Exports the reference's five entry points (src/convGemm.h:20-64) on device
 * pointers plus the extended API of include/convgemm.h.  All device work is
 * reached through the extern "C" launchers of cg_internal.h.
 *
 * Per call: validate -> reference quick-return rules -> weight repack when the
 * kernels need another weight order -> plan (tile configuration + split-K)
 * -> implicit GEMM (per batch chunk when a tensor exceeds the 2 GiB buffer
 * window) -> optional split-K reduction.  Drop-in calls then wait for the
 * stream, like the blocking reference.
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

#define CG_VERSION 10100 /* 1.1.0 */
#define CG_WINDOW (1ll << 31) /* bytes one buffer resource can address */

/* ------------------------------------------------------------------ state */

static _Thread_local convgemm_status_t t_last_status = CONVGEMM_SUCCESS;
static _Thread_local char t_last_msg[256] = "";
static _Thread_local int t_plan_cfg = -1, t_plan_splits = 0, t_plan_kernels = 0;

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static hipStream_t g_stream = NULL;

/* device scratch of the quantizer (amax bits, scale) */
static void *g_qscratch = NULL;

/* buffers handed out by alloc_pack_buffs, with their sizes */
#define CG_MAX_PACKS 64
static struct { void *ptr; size_t bytes; } g_packs[CG_MAX_PACKS];

/* forced plan (tuning hook) */
static int g_force_cfg = -1, g_force_splits = 0;

/* optional planner hook installed by the autotuner / cost model */
static cg_plan_hook_t g_plan_hook = NULL;
static void *g_plan_hook_ctx = NULL;

static int g_fail_allocs = 0;

/* bytes one batch chunk may address (test hook; 0: CG_WINDOW) */
static long long g_chunk_window = 0;
void cg_set_chunk_window(long long bytes) { g_chunk_window = bytes > 0 && bytes < CG_WINDOW ? bytes : 0; }
static int g_dgrad_fwd = 1;
void cg_set_dgrad_forward(int on) { g_dgrad_fwd = on != 0; }

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

int cg_forced_config(int *cfg, int *splits) {
    *cfg = g_force_cfg;
    *splits = g_force_splits;
    return g_force_cfg >= 0;
}

void cg_set_plan_hook(cg_plan_hook_t hook, void *ctx) {
    g_plan_hook = hook;
    g_plan_hook_ctx = ctx;
}

void cg_last_plan(int *cfg, int *splits, int *nkernels) {
    if (cfg) *cfg = t_plan_cfg;
    if (splits) *splits = t_plan_splits;
    if (nkernels) *nkernels = t_plan_kernels;
}

void cg_debug_fail_next_allocs(int n) { g_fail_allocs = n; }

static hipError_t lib_malloc(void **p, size_t bytes) {
    if (g_fail_allocs > 0) {
        g_fail_allocs--;
        *p = NULL;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

static convgemm_status_t hip_fail(hipError_t e, const char *what) {
    return set_error(CONVGEMM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int cg_malloc(void **p, size_t bytes) { return (int)lib_malloc(p, bytes); }
void *cg_get_stream(void) { return (void *)g_stream; }
convgemm_status_t cg_hip_fail(int e, const char *what) { return hip_fail((hipError_t)e, what); }

/* ------------------------------------------------------------- shapes */

int convgemm_out_size(int in, int pad, int dil, int k, int stride) {
    if (stride <= 0) return 0;
    return (in + 2 * pad - dil * (k - 1) - 1) / stride + 1;
}

static int elem_of(convgemm_dtype_t dt) {
    return dt == CONVGEMM_BF16 ? CG_ELEM_BF16 : dt == CONVGEMM_FP8_E4M3 ? CG_ELEM_FP8 : CG_ELEM_F32;
}
static int esize_of(int elem) { return elem == CG_ELEM_F32 ? 4 : elem == CG_ELEM_BF16 ? 2 : 1; }
int cg_esize(int elem) { return esize_of(elem); }

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
    long long K = (long long)d->c * d->kh * d->kw;
    /* one image of X / Y and the whole weight tensor must fit one 2 GiB
     * buffer window (larger batches are split into chunks) */
    long long img_x = (long long)d->h * d->w * d->c * 4, img_y = (long long)*ho * *wo * d->kn * 4;
    if (img_x >= CG_WINDOW || img_y >= CG_WINDOW || K * d->kn * 4 >= CG_WINDOW ||
        (long long)d->b * *ho * *wo >= CG_WINDOW)
        return set_error(CONVGEMM_ERR_UNSUPPORTED, "one image or the weights exceed the 2 GiB addressing window");
    if (d->kn > 65535) return set_error(CONVGEMM_ERR_UNSUPPORTED, "kn > 65535");
    if (d->op == CONVGEMM_DGRAD && (d->kh > 64 || d->kw > 64))
        return set_error(CONVGEMM_ERR_UNSUPPORTED, "backward data supports kernels up to 64x64");
    return CONVGEMM_SUCCESS;
}

/* --------------------------------------------------------- workspace */

/* Size of a buffer handed out by alloc_pack_buffs, or 0 for any other pointer
 * (the library does not know how large a caller's own buffer is, so it never
 * writes to one it did not allocate; such a call uses the internal workspace). */
static size_t pack_bytes_of(const void *p) {
    if (!p) return 0;
    size_t r = 0;
    pthread_mutex_lock(&g_lock);
    for (int i = 0; i < CG_MAX_PACKS; i++)
        if (g_packs[i].ptr == p) r = g_packs[i].bytes;
    pthread_mutex_unlock(&g_lock);
    return r;
}

int cg_stream_capturing(void *stream) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing((hipStream_t)stream, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return cs != hipStreamCaptureStatusNone;
}

/* Grow-only device buffer shared by the calls that bring no workspace.
 * A buffer that is outgrown is retired, never freed: queued work, or a graph
 * captured earlier, may still use it (retired buffers stay allocated until
 * the process ends; with growth by 1.25x their total is at most 4x the live
 * one).  While `stream` is being captured the buffer does not grow: the call
 * fails with CONVGEMM_ERR_WORKSPACE and the caller passes a workspace. */
convgemm_status_t cg_growbuf_get(cg_growbuf *gb, size_t need, void *stream, void **out) {
    pthread_mutex_lock(&g_lock);
    if (need > gb->bytes) {
        if (cg_stream_capturing(stream)) {
            pthread_mutex_unlock(&g_lock);
            return set_error(CONVGEMM_ERR_WORKSPACE,
                             "the stream is being captured and the internal %s (%zu bytes) would have to grow to "
                             "%zu bytes: pass a workspace (convgemm_workspace_size bytes) to calls made under capture",
                             gb->what, gb->bytes, need);
        }
        void *nb = NULL;
        size_t sz = need + (need >> 2);
        if (lib_malloc(&nb, sz) != hipSuccess) {
            pthread_mutex_unlock(&g_lock);
            return set_error(CONVGEMM_ERR_ALLOC, "cannot allocate %zu bytes of %s", sz, gb->what);
        }
        if (gb->ptr) gb->retired++; /* the old buffer stays allocated (see above) */
        gb->ptr = nb;
        gb->bytes = sz;
    }
    *out = gb->ptr;
    pthread_mutex_unlock(&g_lock);
    return CONVGEMM_SUCCESS;
}

static cg_growbuf g_ws = {NULL, 0, 0, "workspace"};

static convgemm_status_t internal_ws(size_t need, hipStream_t stream, void **out) {
    return cg_growbuf_get(&g_ws, need, stream, out);
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

/* Taps of kernel rows (or columns) that reach input positions of residue r
 * modulo stride s: (r + pad - k * dil) divisible by s. */
static int class_taps(int r, int pad, int k, int dil, int s) {
    int n = 0;
    for (int i = 0; i < k; i++)
        if ((r + pad - i * dil) % s == 0) n++;
    return n;
}

/* Backward-data launch order of the residue classes: most taps (the longest
 * workgroups) first, ties in class order; 0 (identity) above 16 classes. */
static unsigned long long dgrad_class_order(const cg_kparams *kp) {
    const int n = kp->vs * kp->hs;
    if (n <= 1 || n > 16) return 0;
    int cls[16], w[16];
    for (int c = 0; c < n; c++) {
        cls[c] = c;
        w[c] = class_taps(c / kp->hs, kp->vp, kp->KH, kp->vd, kp->vs) *
               class_taps(c % kp->hs, kp->hp, kp->KW, kp->hd, kp->hs);
    }
    for (int i = 1; i < n; i++) /* stable insertion sort, descending taps */
        for (int j = i; j > 0 && w[cls[j]] > w[cls[j - 1]]; j--) {
            int t = cls[j];
            cls[j] = cls[j - 1];
            cls[j - 1] = t;
        }
    unsigned long long ord = 0;
    for (int i = 0; i < n; i++) ord |= (unsigned long long)cls[i] << (4 * i);
    return ord;
}

/* Fill the tile-space fields of kp for configuration t. */
/* Halo-tile M tiling: whole output rows per tile when a row fits the BM tile
 * rows, else BM-column pieces of a row. */
static void halo_tiling(const cg_kparams *kp, int BM, int *rows, int *cblk) {
    if (kp->Wo <= BM) {
        *rows = kp->Wo > 0 ? BM / kp->Wo : 1;
        *cblk = 1;
    } else {
        *rows = 1;
        *cblk = cdiv(kp->Wo, BM);
    }
}

/* Multiply-and-shift constants of n / d for 0 <= n < 2^31 (the kernels'
 * FastDiv, igemm_common.hpp): s = 31 + ceil(log2 d), m = ceil(2^s / d). */
static void fastdiv_magic(long long d, unsigned *m, int *s) {
    if (d < 1) d = 1;
    int l = 0;
    while ((1ll << l) < d) l++;
    *s = 31 + l;
    *m = (unsigned)(((1ull << *s) + (unsigned long long)d - 1) / (unsigned long long)d);
}

void cg_fill_tiles(cg_kparams *kp, const cg_tilecfg *t, int splits) {
    int op = kp->op;
    kp->raster = t->raster;
    fastdiv_magic((long long)kp->Ho * kp->Wo, &kp->fdm_howo, &kp->fds_howo);
    fastdiv_magic(kp->Wo, &kp->fdm_wo, &kp->fds_wo);
    fastdiv_magic(kp->C, &kp->fdm_c, &kp->fds_c);
    kp->hrows = kp->hcblk = kp->htpi = kp->hpitch = 0;
    fastdiv_magic(1, &kp->fdm_hp, &kp->fds_hp);
    kp->ncls_h = 1;
    kp->ncls_w = 1;
    kp->cls_order = 0;
    if (op == CG_FWD_NHWC || op == CG_FWD_NCHW) {
        kp->M = (int)kp->P;
        kp->N = kp->KN;
        kp->kchunks = cdiv(kp->C, t->BK);
        kp->KT = kp->taps * kp->kchunks;
        if (t->halo) { /* halo tiles: whole output rows, or pieces of one row; split-K divides the
                          channel chunks */
            halo_tiling(kp, t->BM, &kp->hrows, &kp->hcblk);
            kp->htpi = cdiv(kp->Ho, kp->hrows) * kp->hcblk;
            kp->hpitch = ((kp->hcblk == 1 ? kp->Wo : t->BM) - 1) * kp->hs + (kp->KW - 1) * kp->hd + 1;
            fastdiv_magic(kp->hpitch, &kp->fdm_hp, &kp->fds_hp);
            kp->mtiles = kp->B * kp->htpi;
            kp->ntiles = cdiv(kp->N, t->BN);
            if (splits < 1) splits = 1;
            if (splits > kp->kchunks) splits = kp->kchunks > 0 ? kp->kchunks : 1;
            kp->splits = splits;
            return;
        }
    } else if (op == CG_DGRAD_NHWC || op == CG_DGRAD_NCHW) {
        kp->ncls_h = kp->vs;
        kp->ncls_w = kp->hs;
        kp->cls_order = dgrad_class_order(kp);
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

/* Does the largest halo of the problem fit the configuration?  A halo tile
 * covers `rows` whole output rows, or `cols` <= BM columns of one row, and its
 * halo is ((rows - 1) * vs + (KH - 1) * vd + 1) input rows of
 * ((cols - 1) * hs + (KW - 1) * hd + 1) slots (ht_kernel.hpp); full tiles are
 * the largest.  1x1 kernels gain nothing from a halo and are not served. */
int cg_halo_fits(const cg_kparams *kp, const cg_tilecfg *t) {
    if (!t->halo) return 1;
    if (kp->op != CG_FWD_NHWC || kp->taps <= 1 || kp->Ho <= 0 || kp->Wo <= 0) return 0;
    int rpt, cblk;
    halo_tiling(kp, t->BM, &rpt, &cblk);
    const long long rows = rpt < kp->Ho ? rpt : kp->Ho;
    const long long cols = cblk == 1 ? kp->Wo : t->BM;
    const long long hr = (rows - 1) * kp->vs + (long long)(kp->KH - 1) * kp->vd + 1;
    const long long hc = (cols - 1) * kp->hs + (long long)(kp->KW - 1) * kp->hd + 1;
    return hr * hc <= t->halo;
}

/* A plan the kernels can run for this problem: the configuration serves the
 * operation and operand layout, and a halo tile's halo fits. */
int cg_plan_usable(int elem, int cfg, const cg_kparams *kp) {
    cg_tilecfg t;
    if (cfg < 0 || cg_get_tilecfg(elem, cfg, &t) || !cg_tilecfg_supports(elem, cfg, kp->op, kp->vec)) return 0;
    return cg_halo_fits(kp, &t);
}

size_t cg_partial_bytes(const cg_kparams *kp) {
    if (kp->splits <= 1) return 0;
    return (size_t)kp->splits * kp->ncls_h * kp->ncls_w * (size_t)kp->M * kp->N * sizeof(float);
}

/* Analytic default plan ("no learning"): estimate time as waves of
 * workgroups over the 256 CUs times the per-tile MFMA work, penalise small
 * tiles (operand reuse) and charge the split-K partial traffic; in the spirit
 * of the reference's analytic blocking model (gemm_blis.h:121-148). */
void cg_default_plan(int elem, const cg_kparams *base, int *cfg_out, int *splits_out) {
    const int CUS = 256;
    int op = base->op;
    double best = 1e300;
    int best_cfg = -1, best_splits = 1;
    int n = cg_num_tile_shapes(elem); /* raster order 0: the N panels of an M panel adjacent */
    for (int cfg = 0; cfg < n; cfg++) {
        cg_tilecfg t;
        if (cg_get_tilecfg(elem, cfg, &t) || !cg_tilecfg_supports(elem, cfg, op, base->vec) || t.halo) continue;
        cg_kparams kp = *base;
        cg_fill_tiles(&kp, &t, 1);
        long long tiles = (long long)kp.mtiles * kp.ntiles * kp.ncls_h * kp.ncls_w;
        int lds = t.lp ? (t.BM + t.BN) * 128 * t.lp : (t.BM + t.BN) * (t.BK * esize_of(elem) + 16) * 2;
        int per_cu = 163840 / lds;
        if (per_cu > 4) per_cu = 4;
        if (per_cu < 1) per_cu = 1;
        for (int s = 1; s <= 256; s *= 2) {
            if (s > 1 && kp.KT / s < 2) break;
            if (s > 1) {
                cg_kparams k2 = kp;
                cg_fill_tiles(&k2, &t, s);
                if (cg_partial_bytes(&k2) > (size_t)CG_PACK_BYTES) break;
            }
            long long wgs = tiles * s;
            long long slots = (long long)CUS * per_cu;
            double waves = (double)((wgs + slots - 1) / slots);
            double kwork = (double)((kp.KT + s - 1) / s) * t.BK;
            double cost = waves * (double)t.BM * t.BN * kwork * per_cu;
            cost *= 1.0 + 48.0 / ((double)t.BM * t.BN / (t.BM + t.BN));
            if (s > 1) cost += 8.0 * (double)kp.M * kp.N * s * kp.ncls_h * kp.ncls_w;
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
    int prepacked; /* `in` already holds the weights in the kernels' order */
    int noalloc;   /* the allocation-free entry point: no internal workspace, no planner device work */
} cg_wsets;

/* Weight order the kernels read, and the repack that produces it
 * (arguments of cg_launch_permute3); returns 0 when none is needed. */
/* Backward data of an NHWC convolution with unit strides is the forward
 * convolution of dY with the kernel turned by 180 degrees (taps in reverse
 * order, roles of c and kn swapped): dX[b][ih][iw][c] = sum over (kx, ky, kn)
 * of dY[b][ih + kx*vd - p'v][iw + ky*hd - p'h][kn] * W[c][KH-1-kx][KW-1-ky][kn]
 * with p' = (K - 1) * dil - pad, the gather of src/convGemm.c:86-116 read as
 * a convolution.  It then runs on the forward kernels (pipelined and halo
 * tiles).  A negative p' (padding beyond the kernel's reach) keeps the
 * residue-class kernel. */
static int dgrad_as_forward(const convgemm_desc_t *d) {
    return g_dgrad_fwd && d->op == CONVGEMM_DGRAD && d->layout == CONVGEMM_NHWC && d->vstride == 1 &&
           d->hstride == 1 && (d->kh - 1) * d->vdil >= d->vpad && (d->kw - 1) * d->hdil >= d->hpad;
}

/* Rewrite a prepared backward-data problem (kp->Ho, Wo: dY's extent) into its forward form. */
void cg_forward_form(cg_kparams *kp) {
    const int c = kp->C, kn = kp->KN, h = kp->H, w = kp->W;
    kp->op = CG_FWD_NHWC;
    kp->H = kp->Ho;
    kp->W = kp->Wo;
    kp->C = kn;
    kp->KN = c;
    kp->vp = (kp->KH - 1) * kp->vd - kp->vp;
    kp->hp = (kp->KW - 1) * kp->hd - kp->hp;
    kp->Ho = h;
    kp->Wo = w;
    kp->ffd = 1;
}

int cg_dgrad_forward_form_ok(const cg_kparams *kp) {
    return g_dgrad_fwd && kp->op == CG_DGRAD_NHWC && kp->vs == 1 && kp->hs == 1 && (kp->KH - 1) * kp->vd >= kp->vp &&
           (kp->KW - 1) * kp->hd >= kp->hp;
}

static int weight_repack(const convgemm_desc_t *d, int elem, int *d0, int *d1, int *d2, int *perm) {
    int taps = d->kh * d->kw;
    int K = d->c * taps;
    if (d->op == CONVGEMM_WGRAD) return 0;
    if (dgrad_as_forward(d)) { /* [c][tap][kn] -> [c][taps-1-tap][kn]: the forward form's [kn'][tap][c'] */
        if (taps == 1) return 0;
        *d0 = d->c; *d1 = taps; *d2 = d->kn; *perm = 2;
        return 1;
    }
    if (d->op == CONVGEMM_FWD) {
        if (d->layout == CONVGEMM_NHWC) { /* [c][tap][kn] -> [kn][tap][c] */
            *d0 = d->c; *d1 = taps; *d2 = d->kn; *perm = 0;
        } else { /* [kn][c][tap] -> [kn][tap][c] */
            if (taps == 1) return 0; /* a 1x1 kernel is already in that order */
            *d0 = d->kn; *d1 = d->c; *d2 = taps; *perm = 1;
        }
        return 1;
    }
    if (d->layout == CONVGEMM_NCHW) { /* [kn][K] -> [K][kn] */
        *d0 = d->kn; *d1 = 1; *d2 = K; *perm = 0;
        return 1;
    }
    return 0;
}

/* Should the plan read the caller's weights in place (cg_kparams.braw)?  The
 * pipelined and halo-tile forward NHWC kernels can read [c][kh][kw][kn] rows of whole
 * 16-byte kn chunks from a 16-byte aligned tensor; weights already
 * prepacked, and the backward-data forward form (whose weights need their
 * taps reversed), keep the repacked order.  Reading in place saves the
 * repack launch (about 4.5 us) but the kernel then reads its weight operand
 * with transposing LDS reads: measured on the same plans it ran 2-18%
 * (mean 8%) longer in BF16 and 0-5% (mean 3.5%) in FP32, so it pays below
 * roughly 24 (BF16) or 10 (FP32) GFLOP per call.  In FP8 the best in-place
 * plan of a layer ran 1-22% longer than the best repacked one; with the
 * repack launch counted, reading in place paid or tied up to the largest
 * layer measured, 118 GFLOP.  The halo-tile kernel streams only its weights
 * through the pipeline: in place it ran 0-4% longer in FP32 and 0.1-4.6 us
 * longer per call in BF16 and FP8 at 15 GFLOP (about the repack launch at
 * worst), so its limit is 20 GFLOP in BF16 and FP8. */
static int raw_weights_ok(int elem, const cg_kparams *kp, const cg_tilecfg *t, const void *in) {
    if (kp->op != CG_FWD_NHWC || kp->ffd || !kp->vec || t->lp <= 0 ||
        (kp->KN * esize_of(elem)) % 16 || (uintptr_t)in % 16)
        return 0;
    const double flops = 2.0 * (double)kp->P * kp->C * kp->taps * kp->KN;
    if (t->halo) return flops <= (elem == CG_ELEM_F32 ? 10e9 : 20e9);
    return flops <= (elem == CG_ELEM_FP8 ? 100e9 : elem == CG_ELEM_BF16 ? 24e9 : 10e9);
}

int cg_weight_repack(const convgemm_desc_t *d, int elem, int *d0, int *d1, int *d2, int *perm) {
    return weight_repack(d, elem, d0, d1, d2, perm);
}

/* Test hook: backward-data launch order of the residue classes of a problem
 * (order[i] = the i-th class launched, class = rh * hs + rw) and each class's
 * tap count; returns the number of classes, or 0 when the order is the
 * identity (one class, or more than 16). */
int cg_dgrad_class_order(const convgemm_desc_t *d, int *order, int *taps) {
    cg_kparams kp;
    memset(&kp, 0, sizeof(kp));
    kp.KH = d->kh, kp.KW = d->kw, kp.vp = d->vpad, kp.hp = d->hpad;
    kp.vs = d->vstride, kp.hs = d->hstride, kp.vd = d->vdil, kp.hd = d->hdil;
    const unsigned long long ord = dgrad_class_order(&kp);
    if (!ord) return 0;
    const int n = kp.vs * kp.hs;
    for (int i = 0; i < n; i++) {
        order[i] = (int)((ord >> (4 * i)) & 15);
        taps[i] = class_taps(i / kp.hs, kp.vp, kp.KH, kp.vd, kp.vs) * class_taps(i % kp.hs, kp.hp, kp.KW, kp.hd, kp.hs);
    }
    return n;
}

static int vec_ok(int elem, const convgemm_desc_t *d, const void *a, const void *b) {
    int vw = 16 / esize_of(elem);
    if (((uintptr_t)a % 16) || ((uintptr_t)b % 16)) return 0;
    if (d->op == CONVGEMM_FWD) return d->c % vw == 0;
    if (d->op == CONVGEMM_DGRAD) return d->kn % vw == 0;
    /* weight gradient, NHWC: 16-byte groups of channels of X and of kn of dY */
    if (d->layout == CONVGEMM_NHWC) return d->c % vw == 0 && d->kn % vw == 0;
    return 1;
}

/* Bytes per image of the A operand, the output, and (WGRAD) the dY operand. */
static void image_bytes(const convgemm_desc_t *d, int ho, int wo, int es, long long *img_a, long long *img_o,
                        long long *img_b) {
    *img_o = 0;
    *img_b = 0;
    if (d->op == CONVGEMM_FWD) {
        *img_a = (long long)d->h * d->w * d->c * es;
        *img_o = (long long)ho * wo * d->kn * 4;
    } else if (d->op == CONVGEMM_WGRAD) {
        *img_a = (long long)d->h * d->w * d->c * es;
        *img_b = (long long)ho * wo * d->kn * es;
    } else {
        *img_a = (long long)ho * wo * d->kn * es;
        *img_o = (long long)d->h * d->w * d->c * 4;
    }
}

convgemm_status_t cg_prepare(const convgemm_desc_t *d, convgemm_dtype_t dtype, const void *in, const void *x,
                             const void *wsw, cg_kparams *kp, int *elem_out, int *chunk_out) {
    int ho, wo;
    convgemm_status_t st = validate(d, &ho, &wo);
    if (st != CONVGEMM_SUCCESS) return st;
    if (dtype != CONVGEMM_F32 && dtype != CONVGEMM_BF16 && dtype != CONVGEMM_FP8_E4M3)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "unknown dtype %d", (int)dtype);
    const int elem = elem_of(dtype), es = esize_of(elem);
    if (elem == CG_ELEM_FP8 && d->op != CONVGEMM_FWD)
        return set_error(CONVGEMM_ERR_UNSUPPORTED, "FP8 operands are supported for the forward pass only");
    memset(kp, 0, sizeof(*kp));
    kp->op = opkind(d);
    kp->B = d->b; kp->H = d->h; kp->W = d->w; kp->C = d->c; kp->KN = d->kn; kp->KH = d->kh; kp->KW = d->kw;
    kp->vp = d->vpad; kp->hp = d->hpad; kp->vs = d->vstride; kp->hs = d->hstride; kp->vd = d->vdil;
    kp->hd = d->hdil;
    kp->Ho = ho; kp->Wo = wo; kp->taps = d->kh * d->kw;
    kp->scale = 1.0f;
    int rd0, rd1, rd2, rperm;
    const int repack = weight_repack(d, elem, &rd0, &rd1, &rd2, &rperm);
    /* an internal workspace is 256-byte aligned, so only caller buffers can break vec */
    kp->vec = vec_ok(elem, d, x, repack ? wsw : in);
    if (dgrad_as_forward(d)) cg_forward_form(kp);
    long long img_a, img_o, img_b;
    image_bytes(d, ho, wo, es, &img_a, &img_o, &img_b);
    long long img_max = img_a > img_o ? img_a : img_o;
    if (img_b > img_max) img_max = img_b;
    const long long window = g_chunk_window > 0 ? g_chunk_window : CG_WINDOW;
    int chunk = (int)((window - 1) / (img_max > 0 ? img_max : 1));
    if (chunk > d->b) chunk = d->b;
    if (chunk < 1) chunk = 1;
    kp->B = chunk;
    kp->P = (long long)chunk * kp->Ho * kp->Wo;
    *elem_out = elem;
    *chunk_out = chunk;
    return CONVGEMM_SUCCESS;
}

/* Set while a call plans for a stream that is being captured: the planner
 * hook may then not run device work (the cost model's inference). */
static _Thread_local int t_in_capture;
int cg_planning_in_capture(void) { return t_in_capture; }

/* Plan (cfg, splits) for a prepared problem: forced plan, planner hook, analytic model. */
static int choose_plan(int elem, const cg_kparams *kp, int *cfg, int *splits) {
    *cfg = -1;
    *splits = 1;
    if (g_force_cfg >= 0 && cg_plan_usable(elem, g_force_cfg, kp)) {
        *cfg = g_force_cfg;
        *splits = g_force_splits >= 0 ? g_force_splits : 1; /* 0: stream-K */
    } else if (g_plan_hook) {
        g_plan_hook(g_plan_hook_ctx, elem, kp, cfg, splits);
    }
    if (!cg_plan_usable(elem, *cfg, kp)) cg_default_plan(elem, kp, cfg, splits);
    return *cfg;
}

int cg_choose_plan(int elem, const cg_kparams *kp, int *cfg, int *splits) {
    return choose_plan(elem, kp, cfg, splits);
}

static int num_cus(void) {
    static int n = 0;
    if (!n) {
        int dev = 0, v = 0;
        hipGetDevice(&dev);
        n = (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ? v : 256;
    }
    return n;
}

/* The stream-K schedule of a problem with tiles filled: data-parallel tiles
 * and stream-K workgroups.  Whole rounds of tiles, one per CU, run one tile
 * per workgroup (*dp, a multiple of the CU count); only the remaining
 * T mod CUs tiles are cut along K, evenly over at most one workgroup per CU
 * with at least SK_MIN_STEPS K steps each, so every CU gets T*KT/CUs K steps
 * instead of ceil(T/CUs)*KT, and only the remainder tiles need a fix-up.
 * Returns the stream-K workgroups (0: nothing to cut, or no stream-K kernel
 * for this configuration/operation). */
#define SK_MIN_STEPS 4
int cg_sk_split(int elem, int cfg, const cg_kparams *kp, int *dp) {
    *dp = 0;
    if (cg_igemm_sk_occupancy(elem, cfg, kp) <= 0) return 0;
    const long long cus = num_cus();
    const long long tiles = (long long)kp->mtiles * kp->ntiles;
    const long long rem = tiles % cus;
    *dp = (int)(tiles - rem);
    if (rem == 0 || kp->KT < 1) return 0;
    long long g = rem * kp->KT / SK_MIN_STEPS;
    if (g > cus) g = cus;
    if (g < 1) g = 1;
    return (int)g;
}

int cg_sk_groups(int elem, int cfg, const cg_kparams *kp) {
    int dp;
    return cg_sk_split(elem, cfg, kp, &dp);
}

size_t cg_sk_bytes(int groups, const cg_tilecfg *t) { return (size_t)groups * 2 * t->BM * t->BN * sizeof(float); }

static convgemm_status_t run(const convgemm_desc_t *d, float alpha, float beta, const void *in,
                             const void *x, void *out, const convgemm_epilogue_t *epi,
                             convgemm_dtype_t dtype, const float *scales, cg_wsets *ws,
                             hipStream_t stream) {
    cg_kparams kp;
    int elem, chunk;
    convgemm_status_t st = cg_prepare(d, dtype, in, x, ws->prepacked ? in : ws->ws_w, &kp, &elem, &chunk);
    if (st != CONVGEMM_SUCCESS) return st;
    const int es = esize_of(elem);
    int ho, wo; /* the convolution's output extent (kp's may be that of a forward form) */
    validate(d, &ho, &wo);

    const long long P = (long long)d->b * ho * wo;
    const long long K = (long long)d->c * d->kh * d->kw;
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
    if (elem == CG_ELEM_FP8 && !scales)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "FP8 operands need their per-tensor scales");

    kp.out = (float *)out;
    kp.alpha = alpha;
    kp.beta = beta;
    kp.scale = elem == CG_ELEM_FP8 ? scales[0] * scales[1] : 1.0f;
    if (d->op == CONVGEMM_FWD && epi) {
        kp.bias = epi->bias; kp.mean = epi->mean; kp.inv_std = epi->inv_std;
        kp.gamma = epi->gamma; kp.bnb = epi->beta; kp.relu = epi->relu != 0;
    }

    /* weights in the order the kernels read */
    int rd0 = 0, rd1 = 0, rd2 = 0, rperm = 0;
    int repack = weight_repack(d, elem, &rd0, &rd1, &rd2, &rperm) && !ws->prepacked;

    /* operand roles: A = X (FWD, WGRAD) or dY (DGRAD); B = weights (FWD,
     * DGRAD) or dY (WGRAD) */
    const void *bsrc = in;
    kp.A = x;
    long long img_a, img_o, img_b;
    image_bytes(d, ho, wo, es, &img_a, &img_o, &img_b);

    /* plan on the first chunk's shape */
    int cfg, splits;
    /* under capture, or in the allocation-free entry point, the planner may
     * not run device work (the cost model's inference allocates and copies) */
    t_in_capture = ws->noalloc || cg_stream_capturing(stream);
    choose_plan(elem, &kp, &cfg, &splits);
    t_in_capture = 0;
    cg_tilecfg t;
    if (cfg < 0 || cg_get_tilecfg(elem, cfg, &t))
        return set_error(CONVGEMM_ERR_UNSUPPORTED, "no kernel configuration for this problem");
    cg_fill_tiles(&kp, &t, splits > 0 ? splits : 1);
    /* splits == 0: stream-K (forward and weight gradient) */
    kp.sk = splits == 0 ? cg_sk_split(elem, cfg, &kp, &kp.sk_dp) : 0;
    if (!kp.sk) kp.sk_dp = 0;
    /* a pipelined forward plan reads the caller's weights in place: no repack */
    if (repack && raw_weights_ok(elem, &kp, &t, in)) {
        repack = 0;
        kp.braw = 1;
    }
    const size_t wt_bytes = repack ? (size_t)K * d->kn * es : 0;
    if (repack && ws->ws_w && ws->w_bytes < wt_bytes)
        return set_error(CONVGEMM_ERR_WORKSPACE, "weight workspace %zu < %zu bytes", ws->w_bytes, wt_bytes);

    /* workspace: repacked weights + split-K partials / stream-K pieces */
    size_t pbytes = kp.sk ? cg_sk_bytes(kp.sk, &t) : cg_partial_bytes(&kp);
    if (kp.sk && ws->ws_parts && pbytes > ws->parts_bytes) {
        kp.sk = 0; /* no room for the pieces: plain tiles */
        pbytes = 0;
    }
    while (pbytes && ws->ws_parts && pbytes > ws->parts_bytes && kp.splits > 1) {
        cg_fill_tiles(&kp, &t, kp.splits / 2);
        pbytes = cg_partial_bytes(&kp);
    }
    if ((repack && !ws->ws_w) || (pbytes && !ws->ws_parts)) {
        if (ws->noalloc) /* not reached: the scratch provides both parts */
            return set_error(CONVGEMM_ERR_WORKSPACE, "allocation-free call needs a workspace it does not have");
        size_t wt_al = (wt_bytes + 255) & ~(size_t)255;
        void *base = NULL;
        st = internal_ws(wt_al + pbytes, stream, &base);
        if (st != CONVGEMM_SUCCESS) return st;
        if (repack && !ws->ws_w) { ws->ws_w = base; ws->w_bytes = wt_bytes; }
        if (pbytes && !ws->ws_parts) { ws->ws_parts = (char *)base + wt_al; ws->parts_bytes = pbytes; }
    }
    if (repack) {
        int e = cg_launch_permute3(in, ws->ws_w, rd0, rd1, rd2, rperm, es, stream);
        if (e) return hip_fail((hipError_t)e, "weight repack");
        bsrc = ws->ws_w;
        if ((uintptr_t)bsrc % 16) kp.vec = 0;
        if (!cg_tilecfg_supports(elem, cfg, kp.op, kp.vec))
            return set_error(CONVGEMM_ERR_UNSUPPORTED, "misaligned weight workspace for this configuration");
    }
    kp.ws = pbytes ? (float *)ws->ws_parts : NULL;

    int nkern = repack;
    for (int b0 = 0; b0 < d->b; b0 += chunk) {
        cg_kparams c = kp;
        int nb = d->b - b0 < chunk ? d->b - b0 : chunk;
        c.B = nb;
        c.P = (long long)nb * kp.Ho * kp.Wo;
        cg_fill_tiles(&c, &t, kp.splits);
        if (kp.sk) { /* this chunk's own schedule, within the pieces the workspace holds */
            c.sk = cg_sk_split(elem, cfg, &c, &c.sk_dp);
            if (c.sk > kp.sk) c.sk = kp.sk;
            if (!c.sk) c.sk_dp = 0;
        }
        c.A = (const char *)x + (size_t)b0 * img_a;
        c.a_bytes = (long long)nb * img_a;
        if (d->op == CONVGEMM_WGRAD) {
            c.Bm = (const char *)in + (size_t)b0 * img_b;
            c.b_bytes = (long long)nb * img_b;
            if (b0 > 0) c.beta = 1.0f; /* later chunks accumulate into dW */
        } else {
            c.Bm = bsrc;
            c.b_bytes = (long long)K * d->kn * es;
            c.out = (float *)((char *)out + (size_t)b0 * img_o);
        }
        int e = cg_launch_igemm(elem, cfg, &c, stream);
        if (e) return hip_fail((hipError_t)e, "implicit GEMM launch");
        nkern++;
        if (c.sk > 0) {
            e = cg_launch_sk_fixup(&c, t.BM, t.BN, stream);
            if (e) return hip_fail((hipError_t)e, "stream-K fix-up");
            nkern++;
        } else if (c.splits > 1) {
            e = cg_launch_splitk_reduce(&c, stream);
            if (e) return hip_fail((hipError_t)e, "split-K reduction");
            nkern++;
        }
    }
    t_plan_cfg = cfg;
    t_plan_splits = kp.sk ? 0 : kp.splits;
    t_plan_kernels = nkern;
    return ok();
}

size_t convgemm_packed_weights_size(const convgemm_desc_t *d, convgemm_dtype_t dtype) {
    int ho, wo, d0, d1, d2, perm;
    if (validate(d, &ho, &wo) != CONVGEMM_SUCCESS) return 0;
    const int elem = elem_of(dtype);
    if (!weight_repack(d, elem, &d0, &d1, &d2, &perm)) return 0;
    return (size_t)d->c * d->kh * d->kw * d->kn * esize_of(elem);
}

convgemm_status_t convgemm_pack_weights(const convgemm_desc_t *d, convgemm_dtype_t dtype, const void *in,
                                        void *packed, convgemm_stream_t stream) {
    int ho, wo, d0, d1, d2, perm;
    convgemm_status_t st = validate(d, &ho, &wo);
    if (st != CONVGEMM_SUCCESS) return st;
    if (dtype != CONVGEMM_F32 && dtype != CONVGEMM_BF16 && dtype != CONVGEMM_FP8_E4M3)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "unknown dtype %d", (int)dtype);
    const int elem = elem_of(dtype);
    if (!weight_repack(d, elem, &d0, &d1, &d2, &perm)) return ok(); /* read as given: nothing to pack */
    if (!in || !packed) return set_error(CONVGEMM_ERR_INVALID_VALUE, "NULL weights or packed buffer");
    int e = cg_launch_permute3(in, packed, d0, d1, d2, perm, esize_of(elem), stream);
    return e ? hip_fail((hipError_t)e, "weight repack") : ok();
}

convgemm_status_t convgemm_conv_packed(const convgemm_desc_t *desc, float alpha, float beta, const void *packed,
                                       const void *x, void *out, const convgemm_epilogue_t *epi,
                                       convgemm_dtype_t dtype, const float *scales, void *workspace,
                                       size_t ws_bytes, convgemm_stream_t stream) {
    cg_wsets ws = {0};
    ws.prepacked = 1;
    if (workspace && ws_bytes) {
        ws.ws_parts = workspace;
        ws.parts_bytes = ws_bytes;
    }
    return run(desc, alpha, beta, packed, x, out, epi, dtype, scales, &ws, (hipStream_t)stream);
}

convgemm_status_t convgemm_conv_ex(const convgemm_desc_t *desc, float alpha, float beta, const void *in,
                                   const void *x, void *out, const convgemm_epilogue_t *epi,
                                   convgemm_dtype_t dtype, const float *scales, void *workspace, size_t ws_bytes,
                                   convgemm_stream_t stream) {
    cg_wsets ws = {0};
    if (workspace && ws_bytes) {
        /* repacked weights first (256-byte aligned), partials after */
        int ho, wo;
        convgemm_status_t st = validate(desc, &ho, &wo);
        if (st != CONVGEMM_SUCCESS) return st;
        int d0, d1, d2, perm;
        int elem = elem_of(dtype);
        size_t wt = weight_repack(desc, elem, &d0, &d1, &d2, &perm)
                        ? (size_t)desc->c * desc->kh * desc->kw * desc->kn * esize_of(elem) : 0;
        size_t wt_al = (wt + 255) & ~(size_t)255;
        if (wt_al > ws_bytes)
            return set_error(CONVGEMM_ERR_WORKSPACE, "workspace %zu < %zu bytes needed for weights", ws_bytes, wt_al);
        ws.ws_w = wt ? workspace : NULL;
        ws.w_bytes = wt;
        /* partials after the weights; no room disables split-K */
        ws.ws_parts = (char *)workspace + wt_al;
        ws.parts_bytes = ws_bytes - wt_al;
    }
    return run(desc, alpha, beta, in, x, out, epi, dtype, scales, &ws, (hipStream_t)stream);
}

/* convgemm_conv_ex on the static device scratch (scratch.hip) instead of any
 * workspace: no device allocation, no free, no planner device work (plans
 * come from a forced plan, the tuning database, the cost model's cache of
 * earlier choices, else the analytic planner).  The repacked weights take the
 * front of the scratch; split-K partials get the rest, and plans whose
 * partials do not fit fall back to fewer splits (stream-K to plain tiles).
 * Every call on a device shares that device's scratch: calls that overlap in
 * time on one GPU (other streams, other threads) race.  Nothing in the
 * scratch has to start at zero: every partial a reduction reads is written
 * by the same call first. */
convgemm_status_t convgemm_conv_noalloc(const convgemm_desc_t *desc, float alpha, float beta, const void *in,
                                        const void *x, void *out, const convgemm_epilogue_t *epi,
                                        convgemm_dtype_t dtype, const float *scales, convgemm_stream_t stream) {
    int ho, wo;
    convgemm_status_t st = validate(desc, &ho, &wo);
    if (st != CONVGEMM_SUCCESS) return st;
    if (dtype != CONVGEMM_F32 && dtype != CONVGEMM_BF16 && dtype != CONVGEMM_FP8_E4M3)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "unknown dtype %d", (int)dtype);
    void *scr = NULL;
    size_t sbytes = 0;
    int e = cg_scratch_get(&scr, &sbytes);
    if (e) return hip_fail((hipError_t)e, "device address of the static scratch");
    int d0, d1, d2, perm;
    const int elem = elem_of(dtype);
    const size_t wt = weight_repack(desc, elem, &d0, &d1, &d2, &perm)
                          ? (size_t)desc->c * desc->kh * desc->kw * desc->kn * esize_of(elem) : 0;
    const size_t wt_al = (wt + 255) & ~(size_t)255;
    if (wt_al > sbytes)
        return set_error(CONVGEMM_ERR_WORKSPACE,
                         "allocation-free call: the repacked weights need %zu bytes, more than the %zu-byte static "
                         "scratch; use convgemm_conv_ex with a workspace, or convgemm_conv_packed",
                         wt_al, sbytes);
    cg_wsets ws = {0};
    ws.noalloc = 1;
    ws.ws_w = wt ? scr : NULL;
    ws.w_bytes = wt;
    ws.ws_parts = (char *)scr + wt_al;
    ws.parts_bytes = sbytes - wt_al;
    return run(desc, alpha, beta, in, x, out, epi, dtype, scales, &ws, (hipStream_t)stream);
}

size_t convgemm_noalloc_scratch_bytes(void) { return (size_t)CG_SCRATCH_BYTES; }

convgemm_status_t convgemm_plan(const convgemm_desc_t *d, convgemm_dtype_t dtype, int *cfg, int *splits) {
    cg_kparams kp;
    int elem, chunk;
    if (cfg) *cfg = -1;
    /* plan as for 256-byte-aligned operands */
    convgemm_status_t st = cg_prepare(d, dtype, (void *)256, (void *)256, NULL, &kp, &elem, &chunk);
    if (st != CONVGEMM_SUCCESS) return st;
    int c, s;
    choose_plan(elem, &kp, &c, &s);
    cg_tilecfg t;
    if (c >= 0 && !cg_get_tilecfg(elem, c, &t)) {
        cg_fill_tiles(&kp, &t, s > 0 ? s : 1);
        s = (s == 0 && cg_sk_groups(elem, c, &kp) > 0) ? 0 : kp.splits;
    }
    if (cfg) *cfg = c;
    if (splits) *splits = s;
    return ok();
}

size_t convgemm_workspace_size(const convgemm_desc_t *d, convgemm_dtype_t dtype) {
    int ho, wo;
    if (validate(d, &ho, &wo) != CONVGEMM_SUCCESS) return 0;
    int elem = elem_of(dtype), d0, d1, d2, perm;
    size_t wt = weight_repack(d, elem, &d0, &d1, &d2, &perm)
                    ? (size_t)d->c * d->kh * d->kw * d->kn * esize_of(elem) : 0;
    size_t wt_al = (wt + 255) & ~(size_t)255;
    /* the largest partial buffer any plan may use (plans keep split-K
     * partials and stream-K pieces within CG_PACK_BYTES) */
    cg_kparams kp;
    memset(&kp, 0, sizeof(kp));
    kp.op = opkind(d);
    kp.B = d->b; kp.H = d->h; kp.W = d->w; kp.C = d->c; kp.KN = d->kn; kp.KH = d->kh; kp.KW = d->kw;
    kp.vs = d->vstride; kp.hs = d->hstride;
    kp.vp = d->vpad; kp.hp = d->hpad; kp.vd = d->vdil; kp.hd = d->hdil;
    kp.Ho = ho; kp.Wo = wo; kp.taps = d->kh * d->kw;
    if (dgrad_as_forward(d)) cg_forward_form(&kp);
    kp.P = (long long)d->b * kp.Ho * kp.Wo;
    size_t parts = 0;
    int n = cg_num_tilecfgs(elem);
    for (int cfg = 0; cfg < n; cfg++) {
        cg_tilecfg t;
        if (cg_get_tilecfg(elem, cfg, &t)) continue;
        cg_fill_tiles(&kp, &t, 256);
        size_t b = cg_partial_bytes(&kp);
        if (b > (size_t)CG_PACK_BYTES) b = CG_PACK_BYTES;
        if (b > parts) parts = b;
        cg_fill_tiles(&kp, &t, 1);
        kp.vec = 1;
        b = cg_sk_bytes(cg_sk_groups(elem, cfg, &kp), &t); /* stream-K pieces */
        kp.vec = 0;
        if (b > parts) parts = b;
    }
    return wt_al + parts;
}

/* --------------------------------------------------------- quantization */

convgemm_status_t convgemm_quantize(const float *src, void *dst, size_t n, convgemm_dtype_t dtype,
                                    float *scale_out, convgemm_stream_t stream) {
    if (dtype != CONVGEMM_BF16 && dtype != CONVGEMM_FP8_E4M3)
        return set_error(CONVGEMM_ERR_INVALID_VALUE, "quantize target must be BF16 or FP8_E4M3");
    if (n == 0) {
        if (scale_out) *scale_out = 1.0f;
        return ok();
    }
    if (!src || !dst) return set_error(CONVGEMM_ERR_INVALID_VALUE, "NULL tensor");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == CONVGEMM_BF16) {
        int e = cg_launch_quantize(src, dst, (long long)n, CG_ELEM_BF16, NULL, s);
        if (e) return hip_fail((hipError_t)e, "bf16 conversion");
        if (scale_out) *scale_out = 1.0f;
        return ok();
    }
    pthread_mutex_lock(&g_lock);
    if (!g_qscratch && lib_malloc(&g_qscratch, 256) != hipSuccess) {
        g_qscratch = NULL;
        pthread_mutex_unlock(&g_lock);
        return set_error(CONVGEMM_ERR_ALLOC, "cannot allocate quantizer scratch");
    }
    void *scr = g_qscratch;
    pthread_mutex_unlock(&g_lock);
    int e = cg_launch_quantize(src, dst, (long long)n, CG_ELEM_FP8, scr, s);
    if (e) return hip_fail((hipError_t)e, "fp8 quantization");
    if (scale_out) {
        hipError_t he = hipMemcpyAsync(scale_out, (char *)scr + 4, sizeof(float), hipMemcpyDeviceToHost, s);
        if (he == hipSuccess) he = hipStreamSynchronize(s);
        if (he != hipSuccess) return hip_fail(he, "reading the fp8 scale");
    }
    return ok();
}

/* ------------------------------------------------------------- drop-in */

static void dropin(const convgemm_desc_t *d, float beta, const float *in, const float *x, float *out,
                   const convgemm_epilogue_t *epi, float *ac_pack, float *bc_pack) {
    cg_wsets ws = {0};
    /* Ac_pack holds split-K partials, Bc_pack the repacked weights */
    ws.parts_bytes = pack_bytes_of(ac_pack);
    ws.w_bytes = pack_bytes_of(bc_pack);
    if (ws.parts_bytes) ws.ws_parts = ac_pack;
    if (ws.w_bytes) ws.ws_w = bc_pack;
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
