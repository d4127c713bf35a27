/*
 * batch.c -- many independent convolutions in a few grouped launches
 * (include/convgemm_batch.h).
 *
 * Creation plans every problem as convgemm_conv_ex would (same validation,
 * quick-return rules and planner), then sorts the problems by kernel
 * instance (tile configuration, operation, vector loads) and, inside an
 * instance, by K tiles per workgroup item, longest first.  Each instance is
 * one grouped launch over the tiles of all its problems.  A run is one
 * element-wise launch (weight repacks and beta-only scalings), the grouped
 * implicit GEMMs and one split-K reduction over the problems of every
 * operation.  With CONVGEMM_BATCH_BOUNDED, when the plans name more than
 * CG_BATCH_MAX_GROUPS instances, the problems of the instances with the
 * least work move to the configuration of the largest instance of their
 * operation that serves them until the count fits, so a run is at most
 * CG_BATCH_MAX_GROUPS + 2 kernels.  With CONVGEMM_BATCH_SINGLE the forward
 * NHWC problems with 16-byte channel chunks are planned on the shapes of the
 * single-launch kernel (kernels/single.hip) and run, whatever their shapes,
 * in ONE persistent launch that walks them longest first.  Every
 * problem's item range starts at a multiple of 8 so the kernels can deal each
 * problem's tiles to the 8 XCDs in contiguous runs.  One device allocation
 * holds the problem and job tables, the repacked weights and the split-K
 * partials.
 */
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "convgemm.h"
#include "convgemm_batch.h"
#include "../cg_internal.h"
#include "cg_host.h"

#define AL256(x) (((size_t)(x) + 255) & ~(size_t)255)
#define CG_BATCH_MAX_GROUPS 16

typedef struct {
    cg_kparams kp;
    int cfg;
    long long items; /* workgroup items of the implicit GEMM */
    long long cost;  /* K tiles per item: the ordering key */
    int repack, d0, d1, d2, perm;
    const void *wsrc;
    size_t wbytes, pbytes;
    int single; /* runs in the single-launch group */
    int idx;    /* position in the caller's problem list */
} bprob;

typedef struct {
    int cfg;        /* -1: the single-launch group */
    cg_kparams kp0; /* op and vec of the group */
    cg_group g;
    const int *var; /* single-launch group: each problem's tile shape (device) */
} glaunch;

typedef struct {
    float *out;
    long long n;
    float beta;
} scale_t;

/* A grouped element-wise launch over a device job table. */
typedef struct {
    const cg_ewjob *jobs;
    const int *start;
    int njobs, total, grid;
} ewlaunch;

struct convgemm_batch_s {
    int elem;
    unsigned flags;
    void *dmem;
    int ngl, nsc;
    glaunch *gl;
    int has_red;
    cg_group red;    /* split-K reductions of every operation */
    ewlaunch ew_run; /* per run: repacks (unless static) and beta scalings */
    ewlaunch ew_rep; /* static weights: the repacks, run once at creation */
    scale_t *sc;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int n;       /* problems given at creation */
    int *plans;  /* per given problem: cfg, splits (-1, 0: no GEMM) */
};

static convgemm_status_t errf(convgemm_status_t st, const char *msg) { return cg_set_error(st, msg); }

/* The single-launch problems first, longest first; the others by kernel
 * instance, longest first within one. */
static int cmp_prob(const void *a, const void *b) {
    const bprob *x = (const bprob *)a, *y = (const bprob *)b;
    if (x->single != y->single) return y->single - x->single;
    if (x->single) {
        if (x->cost != y->cost) return x->cost > y->cost ? -1 : 1;
        return x->idx - y->idx;
    }
    if (x->cfg != y->cfg) return x->cfg - y->cfg;
    if (x->kp.op != y->kp.op) return x->kp.op - y->kp.op;
    if (x->kp.vec != y->kp.vec) return x->kp.vec - y->kp.vec;
    if (x->cost != y->cost) return x->cost > y->cost ? -1 : 1;
    return 0;
}

/* The plan of a single-launch problem: a forced plan when the kernel runs
 * its tile shape; otherwise the kernel's shape and raster order (split-K 1)
 * with the least CU time.  Inside one launch that other problems share, a
 * plan costs the time its workgroups hold their slots: its standalone time
 * (tuning database, else the cost model) times the slots it fills,
 * min(workgroups, slots); without timings, its tiles times their MFMA work
 * with the analytic planner's small-tile penalty.  Returns the
 * configuration, or -1 if no shape serves the problem. */
static int plan_single(int elem, const cg_kparams *kp, long long slots) {
    int cfg, splits;
    if (cg_forced_config(&cfg, &splits) && splits <= 1 && cg_plan_usable(elem, cfg, kp) &&
        cg_single_variant(elem, cfg) >= 0)
        return cfg;
    const int tot = cg_num_tilecfgs(elem);
    int *cfgs = (int *)malloc(sizeof(int) * (size_t)(tot > 0 ? tot : 1));
    int *spl = (int *)malloc(sizeof(int) * (size_t)(tot > 0 ? tot : 1));
    float *us = (float *)malloc(sizeof(float) * (size_t)(tot > 0 ? tot : 1));
    int n = 0, r = -1;
    if (cfgs && spl && us) {
        for (int c = 0; c < tot; c++)
            if (cg_single_variant(elem, c) >= 0 && cg_plan_usable(elem, c, kp)) {
                cfgs[n] = c;
                spl[n++] = 1;
            }
        const int timed = cg_plan_times(elem, kp, cfgs, spl, n, us);
        double best = 1e300;
        for (int i = 0; i < n; i++) {
            cg_tilecfg t;
            if (cg_get_tilecfg(elem, cfgs[i], &t)) continue;
            cg_kparams k = *kp;
            cg_fill_tiles(&k, &t, 1);
            const long long w = (long long)k.mtiles * k.ntiles;
            double cost;
            if (timed) {
                if (us[i] < 0.f) continue; /* not timed by the database */
                cost = (double)us[i] * (double)(w < slots ? w : slots);
            } else {
                cost = (double)w * t.BM * t.BN * (double)k.KT * t.BK;
                cost *= 1.0 + 48.0 / ((double)t.BM * t.BN / (t.BM + t.BN));
            }
            if (cost < best) {
                best = cost;
                r = cfgs[i];
            }
        }
    }
    free(cfgs);
    free(spl);
    free(us);
    return r;
}

/* Plan one problem.  *kind: 1 = implicit GEMM (bp filled), 2 = beta scaling
 * only (sc filled, alpha == 0), 0 = no-op. */
static convgemm_status_t plan_one(const convgemm_problem_t *q, convgemm_dtype_t dtype, unsigned flags,
                                  long long slots, bprob *bp, scale_t *sc, int *kind) {
    const convgemm_desc_t *d = &q->desc;
    cg_kparams kp;
    int elem, chunk;
    *kind = 0;
    convgemm_status_t st = cg_prepare(d, dtype, q->in, q->x, NULL, &kp, &elem, &chunk);
    if (st != CONVGEMM_SUCCESS) return st;
    if (chunk < d->b)
        return errf(CONVGEMM_ERR_UNSUPPORTED, "a batch problem exceeds the 2 GiB addressing window");
    const int es = cg_esize(elem);
    /* output pixels of the convolution (kp may describe a backward pass in its forward form) */
    const long long P = (long long)d->b * convgemm_out_size(d->h, d->vpad, d->vdil, d->kh, d->vstride) *
                        convgemm_out_size(d->w, d->hpad, d->hdil, d->kw, d->hstride);
    const long long K = (long long)d->c * d->kh * d->kw;
    long long gm, gn, gk, out_elems;
    if (d->op == CONVGEMM_FWD) {
        gm = d->kn; gn = P; gk = K; out_elems = P * d->kn;
    } else if (d->op == CONVGEMM_WGRAD) {
        gm = d->kn; gn = K; gk = P; out_elems = K * d->kn;
    } else {
        gm = K; gn = P; gk = d->kn; out_elems = (long long)d->b * d->h * d->w * d->c;
    }
    /* the quick-return rules of a single call (src/gemm_blis_B3A2C0_orig.c:45-47, src/quick_gemm.h) */
    if (gm == 0 || gn == 0) return CONVGEMM_SUCCESS;
    if ((q->alpha == 0.0f || gk == 0) && q->beta == 1.0f) return CONVGEMM_SUCCESS;
    if (!q->out) return errf(CONVGEMM_ERR_INVALID_VALUE, "a batch problem has a NULL output pointer");
    if (q->alpha == 0.0f) {
        sc->out = (float *)q->out;
        sc->n = out_elems;
        sc->beta = q->beta;
        *kind = 2;
        return CONVGEMM_SUCCESS;
    }
    if (gk == 0) return CONVGEMM_SUCCESS;
    if (!q->in || !q->x) return errf(CONVGEMM_ERR_INVALID_VALUE, "a batch problem has a NULL input pointer");
    if (q->epi.mean && (!q->epi.inv_std || !q->epi.gamma || !q->epi.beta))
        return errf(CONVGEMM_ERR_INVALID_VALUE, "batch norm needs inv_std, gamma and beta");

    memset(bp, 0, sizeof(*bp));
    kp.out = (float *)q->out;
    kp.alpha = q->alpha;
    kp.beta = q->beta;
    kp.scale = elem == CG_ELEM_FP8 ? q->scales[0] * q->scales[1] : 1.0f;
    if (d->op == CONVGEMM_FWD) {
        kp.bias = q->epi.bias; kp.mean = q->epi.mean; kp.inv_std = q->epi.inv_std;
        kp.gamma = q->epi.gamma; kp.bnb = q->epi.beta; kp.relu = q->epi.relu != 0;
    }
    kp.A = q->x;
    kp.a_bytes = d->op == CONVGEMM_DGRAD ? P * d->kn * es : (long long)d->b * d->h * d->w * d->c * es;
    kp.Bm = q->in;
    kp.b_bytes = d->op == CONVGEMM_WGRAD ? P * d->kn * es : K * d->kn * es;
    bp->repack = cg_weight_repack(d, elem, &bp->d0, &bp->d1, &bp->d2, &bp->perm);
    bp->wsrc = q->in;
    bp->wbytes = bp->repack ? (size_t)(K * d->kn * es) : 0;

    int cfg = -1, splits = 1;
    if ((flags & CONVGEMM_BATCH_SINGLE) && kp.op == CG_FWD_NHWC && kp.vec) cfg = plan_single(elem, &kp, slots);
    if (cfg >= 0) {
        bp->single = 1;
    } else {
        cg_choose_plan(elem, &kp, &cfg, &splits);
        if (splits < 1) splits = 1; /* stream-K plans run as plain tiles in a grouped launch */
    }
    cg_tilecfg t;
    if (cfg < 0 || cg_get_tilecfg(elem, cfg, &t))
        return errf(CONVGEMM_ERR_UNSUPPORTED, "no kernel configuration for a batch problem");
    cg_fill_tiles(&kp, &t, splits);
    bp->kp = kp;
    bp->cfg = cfg;
    bp->pbytes = cg_partial_bytes(&kp);
    bp->items = (long long)kp.mtiles * kp.ntiles * kp.splits * kp.ncls_h * kp.ncls_w;
    /* the ordering key: K tiles per item; single-launch items, whose tile shapes
     * differ, by their MFMA work */
    bp->cost = bp->single ? (long long)kp.KT * t.BM * t.BN : (kp.KT + kp.splits - 1) / kp.splits;
    *kind = 1;
    return CONVGEMM_SUCCESS;
}

static convgemm_status_t issue(convgemm_batch_t bt, hipStream_t s) {
    int e;
    const ewlaunch *w = &bt->ew_run;
    if (w->njobs && (e = cg_launch_ew_grouped(w->jobs, w->start, w->njobs, w->total, w->grid, cg_esize(bt->elem), s)))
        return cg_hip_fail(e, "weight repacks and scalings");
    for (int i = 0; i < bt->ngl; i++) {
        const glaunch *g = &bt->gl[i];
        if (g->cfg < 0 ? (e = cg_launch_single(bt->elem, &g->g, g->var, s))
                       : (e = cg_launch_igemm_grouped(bt->elem, g->cfg, &g->kp0, &g->g, s)))
            return cg_hip_fail(e, "grouped implicit GEMM");
    }
    if (bt->has_red && (e = cg_launch_splitk_reduce_mixed(&bt->red, s)))
        return cg_hip_fail(e, "grouped split-K reduction");
    return CONVGEMM_SUCCESS;
}

static void release(convgemm_batch_t bt) {
    if (!bt) return;
    if (bt->exec) hipGraphExecDestroy(bt->exec);
    if (bt->graph) hipGraphDestroy(bt->graph);
    if (bt->dmem) hipFree(bt->dmem);
    free(bt->gl);
    free(bt->sc);
    free(bt->plans);
    free(bt);
}

static int same_instance(const bprob *x, const bprob *y) {
    if (x->single || y->single) return x->single == y->single;
    return x->cfg == y->cfg && x->kp.op == y->kp.op && x->kp.vec == y->kp.vec;
}

/* Move the problems of the least-work instances to the configuration of the
 * largest instance of their operation and load width that serves all of
 * them, until at most CG_BATCH_MAX_GROUPS instances remain (or no move is
 * possible).  bp is sorted by instance; it is sorted again after each move. */
static void consolidate(int elem, bprob *bp, int np) {
    for (;;) {
        int ng = 0;
        for (int i = 0; i < np; i++)
            if (i == 0 || !same_instance(&bp[i - 1], &bp[i])) ng++;
        if (ng <= CG_BATCH_MAX_GROUPS) return;
        /* instances as [first, end) ranges with their work (items x K tiles) */
        int *first = (int *)malloc(sizeof(int) * (size_t)(ng + 1));
        double *work = (double *)calloc((size_t)ng, sizeof(double));
        char *stuck = (char *)calloc((size_t)ng, 1); /* instances that no sibling serves */
        if (!first || !work || !stuck) {
            free(first);
            free(work);
            free(stuck);
            return;
        }
        int g = -1;
        for (int i = 0; i < np; i++) {
            if (i == 0 || !same_instance(&bp[i - 1], &bp[i])) first[++g] = i;
            work[g] += (double)bp[i].items * (double)bp[i].cost;
        }
        first[ng] = np;
        int moved = 0;
        /* the lightest instance that has a heavier sibling serving all its problems */
        for (int tries = 0; tries < ng && !moved; tries++) {
            int src = -1;
            for (int k = 0; k < ng; k++)
                if (!stuck[k] && (src < 0 || work[k] < work[src])) src = k;
            if (src < 0) break;
            int dst = -1;
            for (int k = 0; k < ng; k++) {
                if (k == src || bp[first[k]].kp.op != bp[first[src]].kp.op ||
                    bp[first[k]].kp.vec != bp[first[src]].kp.vec || work[k] < work[src])
                    continue;
                int ok = 1;
                for (int i = first[src]; i < first[src + 1] && ok; i++)
                    ok = cg_plan_usable(elem, bp[first[k]].cfg, &bp[i].kp);
                if (ok && (dst < 0 || work[k] > work[dst])) dst = k;
            }
            if (dst < 0) {
                stuck[src] = 1; /* cannot move: try the next lightest */
                continue;
            }
            const int cfg = bp[first[dst]].cfg;
            cg_tilecfg t;
            if (cg_get_tilecfg(elem, cfg, &t)) break;
            for (int i = first[src]; i < first[src + 1]; i++) {
                bprob *q = &bp[i];
                cg_fill_tiles(&q->kp, &t, q->kp.splits);
                q->cfg = cfg;
                q->pbytes = cg_partial_bytes(&q->kp);
                q->items = (long long)q->kp.mtiles * q->kp.ntiles * q->kp.splits * q->kp.ncls_h * q->kp.ncls_w;
                q->cost = (q->kp.KT + q->kp.splits - 1) / q->kp.splits;
            }
            moved = 1;
        }
        free(first);
        free(work);
        free(stuck);
        if (!moved) return;
        qsort(bp, (size_t)np, sizeof(bprob), cmp_prob);
    }
}

/* Item ranges (cg_ew_items) of njobs jobs into start[];
 * returns the total, or -1 past 2^31. */
static long long ew_ranges(const cg_ewjob *jobs, int njobs, int *start) {
    long long acc = 0;
    for (int j = 0; j < njobs; j++) {
        start[j] = (int)acc;
        acc += cg_ew_items(&jobs[j]);
        if (acc > INT_MAX) return -1;
    }
    start[njobs] = (int)acc;
    return acc;
}

convgemm_status_t convgemm_batch_create(const convgemm_problem_t *problems, int n, convgemm_dtype_t dtype,
                                        unsigned flags, convgemm_batch_t *batch) {
    if (!batch) return errf(CONVGEMM_ERR_INVALID_VALUE, "batch output pointer is NULL");
    *batch = NULL;
    if (n < 0 || (n > 0 && !problems)) return errf(CONVGEMM_ERR_INVALID_VALUE, "invalid problem list");
    if (flags & ~(unsigned)(CONVGEMM_BATCH_GRAPH | CONVGEMM_BATCH_PERSISTENT | CONVGEMM_BATCH_STATIC_WEIGHTS |
                            CONVGEMM_BATCH_BOUNDED | CONVGEMM_BATCH_SINGLE))
        return errf(CONVGEMM_ERR_INVALID_VALUE, "unknown batch flags");
    if (dtype != CONVGEMM_F32 && dtype != CONVGEMM_BF16 && dtype != CONVGEMM_FP8_E4M3)
        return errf(CONVGEMM_ERR_INVALID_VALUE, "unknown dtype");

    convgemm_batch_t bt = (convgemm_batch_t)calloc(1, sizeof(*bt));
    bprob *bp = (bprob *)calloc((size_t)(n > 0 ? n : 1), sizeof(bprob));
    int *hstart = NULL;
    cg_kparams *htab = NULL;
    cg_ewjob *hjobs = NULL;
    convgemm_status_t st = CONVGEMM_SUCCESS;
    if (!bt || !bp) {
        st = errf(CONVGEMM_ERR_ALLOC, "host allocation failed");
        goto fail;
    }
    bt->elem = dtype == CONVGEMM_BF16 ? CG_ELEM_BF16 : dtype == CONVGEMM_FP8_E4M3 ? CG_ELEM_FP8 : CG_ELEM_F32;
    bt->flags = flags;
    bt->sc = (scale_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(scale_t));
    bt->plans = (int *)malloc(sizeof(int) * 2 * (size_t)(n > 0 ? n : 1));
    bt->n = n;
    if (!bt->sc || !bt->plans) {
        st = errf(CONVGEMM_ERR_ALLOC, "host allocation failed");
        goto fail;
    }

    /* ---- plan every problem ---- */
    int dev = 0, ncu = 0;
    hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
    const int socc = (flags & CONVGEMM_BATCH_SINGLE) ? cg_single_occupancy(bt->elem) : 0;
    const long long slots = (long long)(socc > 0 ? socc : 1) * ncu; /* resident single-launch workgroups */
    int np = 0;
    for (int i = 0; i < n; i++) {
        int kind;
        bt->plans[2 * i] = -1;
        bt->plans[2 * i + 1] = 0;
        st = plan_one(&problems[i], dtype, flags, slots, &bp[np], &bt->sc[bt->nsc], &kind);
        if (st != CONVGEMM_SUCCESS) goto fail;
        if (kind == 1) bp[np++].idx = i;
        else if (kind == 2) bt->nsc++;
    }
    qsort(bp, (size_t)np, sizeof(bprob), cmp_prob);
    int nsingle = 0;
    while (nsingle < np && bp[nsingle].single) nsingle++;
    if (flags & CONVGEMM_BATCH_BOUNDED) consolidate(bt->elem, bp + nsingle, np - nsingle);
    for (int i = 0; i < np; i++) {
        bt->plans[2 * bp[i].idx] = bp[i].cfg;
        bt->plans[2 * bp[i].idx + 1] = bp[i].kp.splits;
    }

    /* ---- groups, item ranges, device layout ---- */
    int ngl = 0, nsplit = 0, nrep = 0;
    for (int i = 0; i < np; i++) {
        if (bp[i].kp.splits > 1) nsplit++;
        if (bp[i].repack) nrep++;
    }
    for (int i = 0; i < np; i++)
        if (i == 0 || !same_instance(&bp[i - 1], &bp[i])) ngl++;
    const int static_w = (flags & CONVGEMM_BATCH_STATIC_WEIGHTS) != 0;
    /* job table: [repacks][scalings]; per run: the repacks unless static, then the scalings */
    const int njobs = nrep + bt->nsc;
    const size_t tab_bytes = AL256(sizeof(cg_kparams) * (size_t)(np + nsplit));
    const size_t job_bytes = AL256(sizeof(cg_ewjob) * (size_t)(njobs > 0 ? njobs : 1));
    const int nstart = np + ngl + nsplit + 1 + (njobs + 1) + (nrep + 1) + nsingle;
    const size_t start_bytes = AL256(sizeof(int) * (size_t)nstart);
    size_t wtot = 0, ptot = 0;
    for (int i = 0; i < np; i++) {
        wtot += AL256(bp[i].wbytes);
        ptot += AL256(bp[i].pbytes);
    }
    const size_t total_bytes = tab_bytes + job_bytes + start_bytes + wtot + ptot;
    if (np > 0 || bt->nsc > 0) {
        int e = cg_malloc(&bt->dmem, total_bytes ? total_bytes : 256);
        if (e) {
            bt->dmem = NULL;
            st = errf(CONVGEMM_ERR_ALLOC, "cannot allocate the batch workspace");
            goto fail;
        }
    }
    char *base = (char *)bt->dmem;
    cg_kparams *dtab = (cg_kparams *)base;
    cg_ewjob *djobs = (cg_ewjob *)(base + tab_bytes);
    int *dstart = (int *)(base + tab_bytes + job_bytes);
    char *wptr = base + tab_bytes + job_bytes + start_bytes;
    char *pptr = wptr + wtot;

    bt->gl = (glaunch *)calloc((size_t)(ngl > 0 ? ngl : 1), sizeof(glaunch));
    htab = (cg_kparams *)calloc((size_t)(np + nsplit > 0 ? np + nsplit : 1), sizeof(cg_kparams));
    hstart = (int *)calloc((size_t)nstart, sizeof(int));
    hjobs = (cg_ewjob *)calloc((size_t)(njobs > 0 ? njobs : 1), sizeof(cg_ewjob));
    if (!bt->gl || !htab || !hstart || !hjobs) {
        st = errf(CONVGEMM_ERR_ALLOC, "host allocation failed");
        goto fail;
    }
    /* workspaces and element-wise jobs */
    int nj = 0;
    for (int i = 0; i < np; i++) {
        if (bp[i].repack) {
            cg_ewjob *j = &hjobs[nj++];
            j->src = bp[i].wsrc;
            j->dst = wptr;
            j->kind = bp[i].perm == 2 ? 3 : bp[i].perm; /* 0, 1: transposes; 3: reversed taps */
            j->d0 = bp[i].d0; j->d1 = bp[i].d1; j->d2 = bp[i].d2;
            j->n = (long long)bp[i].d0 * bp[i].d1 * bp[i].d2;
            bp[i].kp.Bm = wptr;
            wptr += AL256(bp[i].wbytes);
        }
        if (bp[i].pbytes) {
            bp[i].kp.ws = (float *)pptr;
            pptr += AL256(bp[i].pbytes);
        }
        htab[i] = bp[i].kp;
    }
    for (int i = 0; i < bt->nsc; i++) {
        cg_ewjob *j = &hjobs[nj++];
        j->dst = bt->sc[i].out;
        j->kind = 2;
        j->n = bt->sc[i].n;
        j->beta = bt->sc[i].beta;
        if (j->n > INT_MAX) {
            st = errf(CONVGEMM_ERR_UNSUPPORTED, "a batch output has more than 2^31 elements");
            goto fail;
        }
    }
    /* implicit-GEMM groups: item ranges padded to multiples of 8 */
    int so = 0; /* offset into hstart */
    for (int i = 0; i < np;) {
        int j = i;
        while (j < np && same_instance(&bp[j], &bp[i])) j++;
        glaunch *g = &bt->gl[bt->ngl++];
        g->cfg = bp[i].single ? -1 : bp[i].cfg;
        g->kp0 = bp[i].kp;
        long long acc = 0;
        for (int k = i; k < j; k++) {
            hstart[so + (k - i)] = (int)acc;
            acc += (bp[k].items + 7) & ~7LL;
            if (acc > INT_MAX) {
                st = errf(CONVGEMM_ERR_UNSUPPORTED, "a batch group has more than 2^31 workgroup items");
                goto fail;
            }
        }
        hstart[so + (j - i)] = (int)acc;
        g->g.plist = dtab + i;
        g->g.start = dstart + so;
        g->g.nprob = j - i;
        g->g.total = (int)acc;
        g->g.grid = (int)acc;
        if (g->cfg < 0) { /* the single launch: persistent, a multiple of 8 workgroups */
            long long cap = slots & ~7LL;
            if (cap < 8) cap = 8;
            if (cap < g->g.grid) g->g.grid = (int)cap;
        } else if (flags & CONVGEMM_BATCH_PERSISTENT) {
            int occ = cg_igemm_occupancy(bt->elem, g->cfg, &g->kp0);
            long long cap = (long long)(occ > 0 ? occ : 1) * ncu;
            cap &= ~7LL;
            if (cap < 8) cap = 8;
            if (cap < g->g.grid) g->g.grid = (int)cap;
        }
        so += j - i + 1;
        if (g->cfg < 0) { /* each problem's tile shape */
            for (int k = i; k < j; k++) hstart[so + (k - i)] = cg_single_variant(bt->elem, bp[k].cfg);
            g->var = dstart + so;
            so += j - i;
        }
        i = j;
    }
    /* split-K reductions of every operation: one grouped launch */
    int nr = 0;
    for (int i = 0; i < np; i++)
        if (bp[i].kp.splits > 1) htab[np + nr++] = bp[i].kp;
    if (nr > 0) {
        long long acc = 0;
        for (int k = 0; k < nr; k++) {
            hstart[so + k] = (int)acc;
            acc += cg_reduce_items(&htab[np + k]);
            if (acc > INT_MAX) {
                st = errf(CONVGEMM_ERR_UNSUPPORTED, "a batch reduction has more than 2^31 blocks");
                goto fail;
            }
        }
        hstart[so + nr] = (int)acc;
        bt->has_red = 1;
        bt->red.plist = dtab + np;
        bt->red.start = dstart + so;
        bt->red.nprob = nr;
        bt->red.total = (int)acc;
        bt->red.grid = (int)acc;
        if ((flags & CONVGEMM_BATCH_PERSISTENT) && bt->red.grid > 8 * ncu) bt->red.grid = 8 * ncu;
        so += nr + 1;
    }
    /* element-wise launches: per run (repacks unless static, scalings), and
     * the static repacks alone (jobs 0 .. nrep-1, run once below) */
    {
        const int first = static_w ? nrep : 0;
        bt->ew_run.njobs = njobs - first;
        if (bt->ew_run.njobs > 0) {
            const long long tot = ew_ranges(hjobs + first, bt->ew_run.njobs, hstart + so);
            if (tot < 0) {
                st = errf(CONVGEMM_ERR_UNSUPPORTED, "a batch has more than 2^31 element-wise blocks");
                goto fail;
            }
            bt->ew_run.jobs = djobs + first;
            bt->ew_run.start = dstart + so;
            bt->ew_run.total = (int)tot;
            bt->ew_run.grid = (int)(tot < 8LL * ncu ? tot : 8LL * ncu);
            so += bt->ew_run.njobs + 1;
        }
        if (static_w && nrep > 0) {
            const long long tot = ew_ranges(hjobs, nrep, hstart + so);
            if (tot < 0) {
                st = errf(CONVGEMM_ERR_UNSUPPORTED, "a batch has more than 2^31 element-wise blocks");
                goto fail;
            }
            bt->ew_rep.jobs = djobs;
            bt->ew_rep.start = dstart + so;
            bt->ew_rep.njobs = nrep;
            bt->ew_rep.total = (int)tot;
            bt->ew_rep.grid = (int)(tot < 8LL * ncu ? tot : 8LL * ncu);
            so += nrep + 1;
        }
    }
    if (np + nr > 0 || njobs > 0) {
        hipError_t e1 = hipSuccess, e2 = hipSuccess, e3 = hipSuccess;
        if (np + nr > 0) e1 = hipMemcpy(dtab, htab, sizeof(cg_kparams) * (size_t)(np + nr), hipMemcpyHostToDevice);
        if (njobs > 0) e2 = hipMemcpy(djobs, hjobs, sizeof(cg_ewjob) * (size_t)njobs, hipMemcpyHostToDevice);
        e3 = hipMemcpy(dstart, hstart, sizeof(int) * (size_t)(so > 0 ? so : 1), hipMemcpyHostToDevice);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
            st = cg_hip_fail(e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3, "copy of the batch tables");
            goto fail;
        }
    }

    /* ---- static weights: repack now, once ---- */
    if (bt->ew_rep.njobs) {
        const ewlaunch *w = &bt->ew_rep;
        int e = cg_launch_ew_grouped(w->jobs, w->start, w->njobs, w->total, w->grid, cg_esize(bt->elem), NULL);
        hipError_t e2 = e ? hipSuccess : hipStreamSynchronize(NULL);
        if (e || e2 != hipSuccess) {
            st = cg_hip_fail(e ? e : (int)e2, "weight repack");
            goto fail;
        }
    }

    /* ---- graph: capture the run once ---- */
    if (flags & CONVGEMM_BATCH_GRAPH) {
        hipStream_t cs;
        hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
        if (e != hipSuccess) {
            st = cg_hip_fail(e, "capture stream");
            goto fail;
        }
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            st = issue(bt, cs);
            hipError_t e2 = hipStreamEndCapture(cs, &bt->graph);
            if (st == CONVGEMM_SUCCESS && e2 != hipSuccess) st = cg_hip_fail(e2, "graph capture");
            if (st == CONVGEMM_SUCCESS) {
                e2 = hipGraphInstantiate(&bt->exec, bt->graph, NULL, NULL, 0);
                if (e2 != hipSuccess) {
                    bt->exec = NULL;
                    st = cg_hip_fail(e2, "graph instantiation");
                }
            }
        } else {
            st = cg_hip_fail(e, "graph capture");
        }
        hipStreamDestroy(cs);
        if (st != CONVGEMM_SUCCESS) goto fail;
    }
    free(bp);
    free(htab);
    free(hstart);
    free(hjobs);
    *batch = bt;
    return cg_set_error(CONVGEMM_SUCCESS, "");

fail:
    free(bp);
    free(htab);
    free(hstart);
    free(hjobs);
    if (bt && bt->dmem) hipDeviceSynchronize();
    release(bt);
    return st;
}

convgemm_status_t convgemm_batch_run(convgemm_batch_t bt, convgemm_stream_t stream) {
    if (!bt) return errf(CONVGEMM_ERR_INVALID_VALUE, "batch is NULL");
    if (bt->exec) {
        hipError_t e = hipGraphLaunch(bt->exec, (hipStream_t)stream);
        if (e != hipSuccess) return cg_hip_fail(e, "graph launch");
        return cg_set_error(CONVGEMM_SUCCESS, "");
    }
    convgemm_status_t st = issue(bt, (hipStream_t)stream);
    if (st != CONVGEMM_SUCCESS) return st;
    return cg_set_error(CONVGEMM_SUCCESS, "");
}

convgemm_status_t convgemm_batch_info(convgemm_batch_t bt, int *groups, int *kernels_per_run, int *graph) {
    if (!bt) return errf(CONVGEMM_ERR_INVALID_VALUE, "batch is NULL");
    if (groups) *groups = bt->ngl;
    if (kernels_per_run) *kernels_per_run = (bt->ew_run.njobs > 0) + bt->ngl + bt->has_red;
    if (graph) *graph = bt->exec != NULL;
    return cg_set_error(CONVGEMM_SUCCESS, "");
}

convgemm_status_t convgemm_batch_plan(convgemm_batch_t bt, int i, int *cfg, int *splits) {
    if (!bt) return errf(CONVGEMM_ERR_INVALID_VALUE, "batch is NULL");
    if (i < 0 || i >= bt->n) return errf(CONVGEMM_ERR_INVALID_VALUE, "problem index out of range");
    if (cfg) *cfg = bt->plans[2 * i];
    if (splits) *splits = bt->plans[2 * i + 1];
    return cg_set_error(CONVGEMM_SUCCESS, "");
}

void convgemm_batch_destroy(convgemm_batch_t bt) {
    if (!bt) return;
    hipDeviceSynchronize();
    release(bt);
}
