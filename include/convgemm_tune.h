/*
 * convgemm_tune.h -- tile-configuration selection for the implicit-GEMM
 * kernels: an exhaustive autotuner with a timing database, and a learned
 * cost model (a small MLP trained and evaluated by the library's own HIP
 * kernels) that predicts the configuration for shapes never tuned.
 *
 * The reference picks its blocking (MC/NC/KC) from the BLIS context and keeps
 * an analytical cache model disabled (src/gemm_blis.h:121-188); these calls
 * have no direct counterpart there -- they replace that blocking model.
 */
#ifndef CONVGEMM_AMD_CONVGEMM_TUNE_H
#define CONVGEMM_AMD_CONVGEMM_TUNE_H

#include "convgemm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Planner used by every convolution call:
 *   CONVGEMM_PLAN_ANALYTIC: the analytic model only;
 *   CONVGEMM_PLAN_TUNED:    tuning database, else analytic;
 *   CONVGEMM_PLAN_LEARNED:  tuning database, else the cost model, else analytic;
 *   CONVGEMM_PLAN_MODEL:    the cost model, else analytic (ignores the database).
 * At start-up the library loads a cost model from the file CONVGEMM_COSTMODEL
 * names ("" or "none": none), else from data/mi355x_costmodel.mlp next to the
 * library; the default planner is CONVGEMM_PLAN_LEARNED when one loaded,
 * CONVGEMM_PLAN_TUNED otherwise.  The model runs on the device, once per new
 * problem shape (its plans are cached); a call on a stream that is being
 * captured only reuses cached plans and otherwise plans without the model. */
typedef enum {
    CONVGEMM_PLAN_ANALYTIC = 0,
    CONVGEMM_PLAN_TUNED = 1,
    CONVGEMM_PLAN_LEARNED = 2,
    CONVGEMM_PLAN_MODEL = 3
} convgemm_planner_t;

void convgemm_set_planner(convgemm_planner_t mode);
convgemm_planner_t convgemm_get_planner(void);

/* Time every supported (tile configuration, split-K) pair on this problem with
 * the caller's buffers (the output is overwritten; backward data accumulates
 * into it), record all timings in the tuning database and make the fastest
 * the plan for this problem.  reps >= 1 timed launches per candidate (median).
 * Any of best_cfg / best_splits / best_us may be NULL. */
convgemm_status_t convgemm_autotune(const convgemm_desc_t *desc, convgemm_dtype_t dtype, const float *scales,
                                    const void *in, const void *x, void *out, int reps, int *best_cfg,
                                    int *best_splits, float *best_us, convgemm_stream_t stream);

/* Timing database: problems tuned so far, persistence, reset. */
int convgemm_tune_count(void);
convgemm_status_t convgemm_tune_save(const char *path);
convgemm_status_t convgemm_tune_load(const char *path);
void convgemm_tune_clear(void);

/* Plan the library would use for this problem now: tile configuration and
 * split-K factor (cfg < 0 if the problem is invalid). */
convgemm_status_t convgemm_plan(const convgemm_desc_t *desc, convgemm_dtype_t dtype, int *cfg, int *splits);

/* Learned cost model: an MLP (features of the problem and the candidate ->
 * predicted log time) trained on the tuning database by HIP kernels
 * (forward, backward, Adam).  Returns the final training loss in *loss. */
convgemm_status_t convgemm_costmodel_train(int epochs, float lr, unsigned seed, float *loss,
                                           convgemm_stream_t stream);
/* Optional progress report for convgemm_costmodel_train: fn(epoch, standardized
 * training loss, ctx) every `every` epochs (fn NULL or every <= 0: none).  Each
 * report synchronizes the training stream once. */
typedef void (*convgemm_train_progress_t)(int epoch, float loss, void *ctx);
void convgemm_costmodel_set_progress(convgemm_train_progress_t fn, int every, void *ctx);
convgemm_status_t convgemm_costmodel_save(const char *path);
convgemm_status_t convgemm_costmodel_load(const char *path);
/* Upper bound on the candidate plans (configuration, split) of one problem:
 * arrays of this length hold every candidate of any problem. */
#define CONVGEMM_MAX_CANDIDATES 4096

/* Predicted time (microseconds) of every candidate of a problem; out[i] is
 * for (cfgs[i], splits[i]).  n_out: capacity in, candidates written out. */
convgemm_status_t convgemm_costmodel_predict(const convgemm_desc_t *desc, convgemm_dtype_t dtype, int *cfgs,
                                             int *splits, float *us, int *n_out);

#ifdef __cplusplus
}
#endif

#endif
