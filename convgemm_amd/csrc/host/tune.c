/*
 * tune.c -- exhaustive autotuner, timing database and the planner hook.
 *
 * The reference takes its blocking (MC/NC/KC) from the BLIS context and keeps
 * an analytical cache model disabled (src/gemm_blis.h:121-188).  Here the
 * "blocking" is a tile configuration plus a split-K factor; this file times
 * every candidate on the device, keeps every timing (the cost model's
 * training data) and serves the fastest as the plan for the tuned shape.
 */
#define _GNU_SOURCE /* dladdr */
#include <dlfcn.h>
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

#define KEYLEN 16
#define MAXCAND CONVGEMM_MAX_CANDIDATES

typedef struct {
    int key[KEYLEN]; /* elem, op, vec, B, H, W, C, KN, KH, KW, vp, hp, vs, hs, vd, hd */
    int best_cfg, best_splits;
    float best_us;
    int n;
    int *cfg, *splits;
    float *us;
    int used;
} tentry;

static pthread_mutex_t t_lock = PTHREAD_MUTEX_INITIALIZER;
static tentry *t_tab = NULL; /* open addressing */
static int t_cap = 0, t_count = 0;
static convgemm_planner_t t_mode = CONVGEMM_PLAN_TUNED;

/* The key of a problem is the call's own description.  The forward form of a
 * backward-data problem keeps the backward-data key: its timings include the
 * read of dX (dX += ...) and must not mix with those of the forward problem
 * that has the same GEMM. */
void cg_key_of(int elem, const cg_kparams *kp, int *k) {
    k[0] = elem; k[1] = kp->op; k[2] = kp->vec; k[3] = kp->B; k[4] = kp->H; k[5] = kp->W; k[6] = kp->C;
    k[7] = kp->KN; k[8] = kp->KH; k[9] = kp->KW; k[10] = kp->vp; k[11] = kp->hp; k[12] = kp->vs;
    k[13] = kp->hs; k[14] = kp->vd; k[15] = kp->hd;
    if (kp->ffd) { /* undo cg_forward_form */
        k[1] = CG_DGRAD_NHWC; k[4] = kp->Ho; k[5] = kp->Wo; k[6] = kp->KN; k[7] = kp->C;
        k[10] = (kp->KH - 1) * kp->vd - kp->vp;
        k[11] = (kp->KW - 1) * kp->hd - kp->hp;
    }
}

static uint64_t key_hash(const int *k) {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < KEYLEN; i++) {
        h ^= (uint32_t)k[i];
        h *= 1099511628211ull;
    }
    return h;
}

static tentry *find_slot(tentry *tab, int cap, const int *k) {
    uint64_t i = key_hash(k) & (uint64_t)(cap - 1);
    for (;;) {
        tentry *e = &tab[i];
        if (!e->used || !memcmp(e->key, k, sizeof(int) * KEYLEN)) return e;
        i = (i + 1) & (uint64_t)(cap - 1);
    }
}

static int grow(void) {
    int ncap = t_cap ? t_cap * 2 : 256;
    tentry *nt = calloc((size_t)ncap, sizeof(tentry));
    if (!nt) return -1;
    for (int i = 0; i < t_cap; i++)
        if (t_tab[i].used) *find_slot(nt, ncap, t_tab[i].key) = t_tab[i];
    free(t_tab);
    t_tab = nt;
    t_cap = ncap;
    return 0;
}

/* Store all timings of one problem (replaces an older record). Caller holds t_lock. */
static int db_put(const int *k, int n, const int *cfg, const int *splits, const float *us) {
    if ((t_count + 1) * 2 > t_cap && grow()) return -1;
    tentry *e = find_slot(t_tab, t_cap, k);
    if (e->used) {
        free(e->cfg); free(e->splits); free(e->us);
    } else {
        t_count++;
    }
    memset(e, 0, sizeof(*e));
    memcpy(e->key, k, sizeof(int) * KEYLEN);
    e->used = 1;
    e->n = n;
    e->cfg = malloc(sizeof(int) * (n ? n : 1));
    e->splits = malloc(sizeof(int) * (n ? n : 1));
    e->us = malloc(sizeof(float) * (n ? n : 1));
    e->best_cfg = -1;
    e->best_us = INFINITY;
    for (int i = 0; i < n; i++) {
        e->cfg[i] = cfg[i];
        e->splits[i] = splits[i];
        e->us[i] = us[i];
        if (us[i] < e->best_us) {
            e->best_us = us[i];
            e->best_cfg = cfg[i];
            e->best_splits = splits[i];
        }
    }
    return 0;
}

int convgemm_tune_count(void) {
    pthread_mutex_lock(&t_lock);
    int n = t_count;
    pthread_mutex_unlock(&t_lock);
    return n;
}

void convgemm_tune_clear(void) {
    pthread_mutex_lock(&t_lock);
    for (int i = 0; i < t_cap; i++)
        if (t_tab[i].used) {
            free(t_tab[i].cfg); free(t_tab[i].splits); free(t_tab[i].us);
        }
    free(t_tab);
    t_tab = NULL;
    t_cap = t_count = 0;
    pthread_mutex_unlock(&t_lock);
}

/* Iterate the database (for the cost model's training set). Returns the
 * number of records; fn is called with the lock held. */
int cg_tune_foreach(void (*fn)(void *ctx, const int *key, int n, const int *cfg, const int *splits,
                               const float *us),
                    void *ctx) {
    pthread_mutex_lock(&t_lock);
    int n = 0;
    for (int i = 0; i < t_cap; i++)
        if (t_tab[i].used) {
            fn(ctx, t_tab[i].key, t_tab[i].n, t_tab[i].cfg, t_tab[i].splits, t_tab[i].us);
            n++;
        }
    pthread_mutex_unlock(&t_lock);
    return n;
}

/* ------------------------------------------------------------ candidates */

/* Split-K factors tried for every configuration. */
static const int S[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256};

/* Test hook: the split-K factors (count returned, up to max written). */
int cg_split_factors(int *out, int max) {
    int n = (int)(sizeof(S) / sizeof(S[0]));
    for (int i = 0; i < n && i < max; i++) out[i] = S[i];
    return n;
}

int cg_candidates(int elem, const cg_kparams *base, int *cfgs, int *splits, int max) {
    int n = 0, ncfg = cg_num_tilecfgs(elem);
    for (int c = 0; c < ncfg; c++) {
        cg_tilecfg t;
        if (cg_get_tilecfg(elem, c, &t) || !cg_plan_usable(elem, c, base)) continue;
        cg_kparams kp = *base;
        cg_fill_tiles(&kp, &t, 1);
        /* a raster order that walks the tiles as order 0 does (one M or one N
         * panel), or as order 1 does (order 2 with at most one band), adds
         * nothing */
        if (t.raster > 0 && (kp.mtiles == 1 || kp.ntiles == 1)) continue;
        if (t.raster == 2 && kp.mtiles <= 8) continue;
        for (unsigned si = 0; si < sizeof(S) / sizeof(S[0]); si++) {
            int s = S[si];
            if (s > 1 && kp.KT / s < 2) break;
            if (t.halo && s > kp.kchunks) break; /* halo tiles split whole channel chunks */
            cg_kparams k2 = kp;
            cg_fill_tiles(&k2, &t, s);
            if (cg_partial_bytes(&k2) > ((size_t)CG_PACK_BYTES)) break;
            if (n < max) {
                cfgs[n] = c;
                splits[n] = s;
                n++;
            }
        }
        /* stream-K (splits 0) when the tiles leave a partial round over the CUs */
        int g = cg_sk_groups(elem, c, &kp);
        if (g > 0 && kp.KT >= 2 && cg_sk_bytes(g, &t) <= (size_t)CG_PACK_BYTES && n < max) {
            cfgs[n] = c;
            splits[n] = 0;
            n++;
        }
    }
    return n;
}

/* --------------------------------------------------------------- planner */

int cg_costmodel_choose(int elem, const cg_kparams *kp, int *cfg, int *splits, int allow_device);
int cg_planning_in_capture(void);

static void plan_hook(void *ctx, int elem, const cg_kparams *kp, int *cfg, int *splits) {
    (void)ctx;
    if (t_mode == CONVGEMM_PLAN_ANALYTIC) return;
    if (t_mode != CONVGEMM_PLAN_MODEL) {
        int k[KEYLEN];
        cg_key_of(elem, kp, k);
        pthread_mutex_lock(&t_lock);
        if (t_cap) {
            tentry *e = find_slot(t_tab, t_cap, k);
            if (e->used && e->best_cfg >= 0) {
                *cfg = e->best_cfg;
                *splits = e->best_splits;
                pthread_mutex_unlock(&t_lock);
                return;
            }
        }
        pthread_mutex_unlock(&t_lock);
    }
    /* under stream capture only plans the model already made (no device inference) */
    if (t_mode == CONVGEMM_PLAN_LEARNED || t_mode == CONVGEMM_PLAN_MODEL)
        cg_costmodel_choose(elem, kp, cfg, splits, !cg_planning_in_capture());
}

int cg_costmodel_times(int elem, const cg_kparams *kp, const int *cfgs, const int *splits, int n, float *us);

/* Times of n candidates (cfgs[i], splits[i]) of one problem, from the
 * evidence the planner would use: the tuning database's recorded timings when
 * it holds the problem (a candidate it did not time gets a negative time),
 * else the cost model's predictions.  Returns 1 with us[] filled, 0 when
 * neither is available (or the stream is being captured and the model would
 * have to run). */
int cg_plan_times(int elem, const cg_kparams *kp, const int *cfgs, const int *splits, int n, float *us) {
    if (n <= 0) return 0;
    if (t_mode != CONVGEMM_PLAN_ANALYTIC && t_mode != CONVGEMM_PLAN_MODEL) {
        int k[KEYLEN], hit = 0;
        cg_key_of(elem, kp, k);
        pthread_mutex_lock(&t_lock);
        if (t_cap) {
            const tentry *e = find_slot(t_tab, t_cap, k);
            if (e->used) {
                for (int i = 0; i < n; i++) {
                    us[i] = -1.f;
                    for (int j = 0; j < e->n; j++)
                        if (e->cfg[j] == cfgs[i] && e->splits[j] == splits[i]) us[i] = e->us[j];
                    hit |= us[i] >= 0.f;
                }
            }
        }
        pthread_mutex_unlock(&t_lock);
        if (hit) return 1;
    }
    if ((t_mode == CONVGEMM_PLAN_LEARNED || t_mode == CONVGEMM_PLAN_MODEL) && !cg_planning_in_capture())
        return cg_costmodel_times(elem, kp, cfgs, splits, n, us) == 0;
    return 0;
}

void convgemm_set_planner(convgemm_planner_t mode) {
    t_mode = mode;
    cg_set_plan_hook(plan_hook, NULL);
}

convgemm_planner_t convgemm_get_planner(void) { return t_mode; }

/* The cost model loaded at start-up: CONVGEMM_COSTMODEL names a file ("" or
 * "none": no model), else data/mi355x_costmodel.mlp next to this library.
 * With a model the default planner is CONVGEMM_PLAN_LEARNED. */
static void load_default_model(void) {
    const char *env = getenv("CONVGEMM_COSTMODEL");
    char path[4096];
    if (env) {
        if (!*env || !strcmp(env, "none")) return;
        snprintf(path, sizeof(path), "%s", env);
    } else {
        Dl_info info;
        if (!dladdr((void *)&load_default_model, &info) || !info.dli_fname) return;
        snprintf(path, sizeof(path), "%s", info.dli_fname);
        char *slash = strrchr(path, '/');
        size_t dir = slash ? (size_t)(slash - path + 1) : 0;
        snprintf(path + dir, sizeof(path) - dir, "data/mi355x_costmodel.mlp");
    }
    if (convgemm_costmodel_load(path) == CONVGEMM_SUCCESS) t_mode = CONVGEMM_PLAN_LEARNED;
    cg_set_error(CONVGEMM_SUCCESS, ""); /* a missing file is not the caller's error */
}

__attribute__((constructor)) static void tune_init(void) {
    cg_set_plan_hook(plan_hook, NULL);
    load_default_model();
    const char *db = getenv("CONVGEMM_TUNING_DB");
    if (db && *db) convgemm_tune_load(db);
}

/* ------------------------------------------------------------- autotune */

static int cmpf(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return x < y ? -1 : x > y;
}

convgemm_status_t convgemm_autotune(const convgemm_desc_t *desc, convgemm_dtype_t dtype, const float *scales,
                                    const void *in, const void *x, void *out, int reps, int *best_cfg,
                                    int *best_splits, float *best_us, convgemm_stream_t stream) {
    cg_kparams kp;
    int elem, chunk;
    convgemm_status_t st = cg_prepare(desc, dtype, in, x, NULL, &kp, &elem, &chunk);
    if (st != CONVGEMM_SUCCESS) return st;
    if (reps < 1) reps = 1;
    if (reps > 64) reps = 64;
    int cfgs[MAXCAND], spl[MAXCAND];
    float us[MAXCAND];
    int n = cg_candidates(elem, &kp, cfgs, spl, MAXCAND);
    if (n == 0) return cg_set_error(CONVGEMM_ERR_UNSUPPORTED, "no kernel configuration for this problem");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return cg_set_error(CONVGEMM_ERR_HIP, "hipEventCreate failed");
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return cg_set_error(CONVGEMM_ERR_HIP, "hipEventCreate failed");
    }
    const float beta = desc->op == CONVGEMM_DGRAD ? 1.0f : 0.0f;
    float samples[64];
    convgemm_status_t fail = CONVGEMM_SUCCESS;
    for (int i = 0; i < n && fail == CONVGEMM_SUCCESS; i++) {
        cg_force_config(cfgs[i], spl[i]);
        us[i] = INFINITY;
        st = convgemm_conv_ex(desc, 1.0f, beta, in, x, out, NULL, dtype, scales, NULL, 0, stream);  // warm-up
        if (st != CONVGEMM_SUCCESS) continue;
        for (int r = 0; r < reps; r++) {
            hipEventRecord(e0, s);
            st = convgemm_conv_ex(desc, 1.0f, beta, in, x, out, NULL, dtype, scales, NULL, 0, stream);
            hipEventRecord(e1, s);
            hipError_t he = hipEventSynchronize(e1);
            float ms = 0.f;
            if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
            if (he != hipSuccess) {
                fail = cg_set_error(CONVGEMM_ERR_HIP, hipGetErrorString(he));
                break;
            }
            samples[r] = st == CONVGEMM_SUCCESS ? ms * 1000.0f : INFINITY;
        }
        if (fail != CONVGEMM_SUCCESS) break;
        qsort(samples, (size_t)reps, sizeof(float), cmpf);
        us[i] = samples[reps / 2];
    }
    cg_force_config(-1, 0);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (fail != CONVGEMM_SUCCESS) return fail;
    int k[KEYLEN];
    cg_key_of(elem, &kp, k);
    int bi = 0;
    for (int i = 1; i < n; i++)
        if (us[i] < us[bi]) bi = i;
    pthread_mutex_lock(&t_lock);
    int rc = db_put(k, n, cfgs, spl, us);
    pthread_mutex_unlock(&t_lock);
    if (rc) return cg_set_error(CONVGEMM_ERR_ALLOC, "tuning database allocation failed");
    if (best_cfg) *best_cfg = cfgs[bi];
    if (best_splits) *best_splits = spl[bi];
    if (best_us) *best_us = us[bi];
    return cg_set_error(CONVGEMM_SUCCESS, "");
}

/* All recorded timings of one problem (prepared as for aligned operands).
 * Returns the number of candidates written (0 if the problem was not tuned). */
int cg_tune_get(const convgemm_desc_t *desc, convgemm_dtype_t dtype, int *cfgs, int *splits, float *us, int max) {
    cg_kparams kp;
    int elem, chunk;
    if (cg_prepare(desc, dtype, (void *)256, (void *)256, NULL, &kp, &elem, &chunk) != CONVGEMM_SUCCESS) return 0;
    int k[KEYLEN];
    cg_key_of(elem, &kp, k);
    int n = 0;
    pthread_mutex_lock(&t_lock);
    if (t_cap) {
        tentry *e = find_slot(t_tab, t_cap, k);
        if (e->used)
            for (int i = 0; i < e->n && n < max; i++, n++) {
                cfgs[n] = e->cfg[i];
                splits[n] = e->splits[i];
                us[n] = e->us[i];
            }
    }
    pthread_mutex_unlock(&t_lock);
    return n;
}

/* ------------------------------------------------------------ persistence */

/* Text format, one problem per line:
 *   16 key ints, candidate count n, then n triples "cfg splits microseconds"
 * after a header that records the tile shapes of each operand type.  A
 * configuration is tile shape base + raster order r * (number of shapes), and
 * tables only grow by appending shapes, so the ids of a database written with
 * fewer shapes are renumbered on load: base + r * (shapes now). */
#define TILE_SHAPES_TAG "# tile shapes (f32 bf16 fp8):"
convgemm_status_t convgemm_tune_save(const char *path) {
    FILE *f = fopen(path, "w");
    if (!f) return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "cannot open tuning database for writing");
    fprintf(f, "# convgemm tuning database v1: elem op vec B H W C KN KH KW vp hp vs hs vd hd | n (cfg splits us)*n\n");
    fprintf(f, TILE_SHAPES_TAG " %d %d %d\n", cg_num_tile_shapes(CG_ELEM_F32), cg_num_tile_shapes(CG_ELEM_BF16),
            cg_num_tile_shapes(CG_ELEM_FP8));
    pthread_mutex_lock(&t_lock);
    for (int i = 0; i < t_cap; i++) {
        tentry *e = &t_tab[i];
        if (!e->used) continue;
        for (int j = 0; j < KEYLEN; j++) fprintf(f, "%d ", e->key[j]);
        fprintf(f, "%d", e->n);
        for (int j = 0; j < e->n; j++) fprintf(f, " %d %d %.3f", e->cfg[j], e->splits[j], isfinite(e->us[j]) ? e->us[j] : 1e30f);
        fprintf(f, "\n");
    }
    pthread_mutex_unlock(&t_lock);
    int bad = ferror(f);
    fclose(f);
    return bad ? cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "write error") : cg_set_error(CONVGEMM_SUCCESS, "");
}

convgemm_status_t convgemm_tune_load(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "cannot open tuning database");
    /* a line holds at most CONVGEMM_MAX_CANDIDATES triples of at most 48 characters */
    const int line_max = 48 * CONVGEMM_MAX_CANDIDATES + 256;
    char *line = malloc(line_max);
    if (!line) {
        fclose(f);
        return cg_set_error(CONVGEMM_ERR_ALLOC, "tuning database: out of host memory");
    }
    int loaded = 0;
    int old_ns[3] = {0, 0, 0}; /* tile shapes of the writer, 0: not recorded (ids taken as they are) */
    while (fgets(line, line_max, f)) {
        if (strncmp(line, TILE_SHAPES_TAG, strlen(TILE_SHAPES_TAG)) == 0) {
            if (sscanf(line + strlen(TILE_SHAPES_TAG), "%d %d %d", &old_ns[0], &old_ns[1], &old_ns[2]) != 3)
                old_ns[0] = old_ns[1] = old_ns[2] = 0;
            continue;
        }
        if (line[0] == '#' || line[0] == '\n') continue;
        int k[KEYLEN], n, pos = 0, used;
        char *p = line;
        int okk = 1;
        for (int j = 0; j < KEYLEN && okk; j++) {
            if (sscanf(p, "%d%n", &k[j], &used) != 1) okk = 0;
            p += used;
        }
        if (!okk || sscanf(p, "%d%n", &n, &used) != 1 || n < 0 || n > MAXCAND) continue;
        p += used;
        int cfg[MAXCAND], spl[MAXCAND];
        float us[MAXCAND];
        for (pos = 0; pos < n; pos++) {
            if (sscanf(p, "%d %d %f%n", &cfg[pos], &spl[pos], &us[pos], &used) != 3) break;
            p += used;
        }
        if (pos != n) continue;
        const int e = k[0];
        if (e >= 0 && e < 3 && old_ns[e] > 0) {
            const int ns = cg_num_tile_shapes(e);
            int kept = 0;
            for (int j = 0; j < n; j++) {
                const int base = cfg[j] % old_ns[e], r = cfg[j] / old_ns[e];
                if (base >= ns || r >= CG_NUM_RASTERS) continue; /* a shape this build does not have */
                cfg[kept] = base + r * ns;
                spl[kept] = spl[j];
                us[kept] = us[j];
                kept++;
            }
            n = kept;
        }
        pthread_mutex_lock(&t_lock);
        db_put(k, n, cfg, spl, us);
        pthread_mutex_unlock(&t_lock);
        loaded++;
    }
    fclose(f);
    free(line);
    (void)loaded;
    return cg_set_error(CONVGEMM_SUCCESS, "");
}
