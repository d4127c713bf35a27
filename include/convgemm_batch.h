/*
 * convgemm_batch.h -- many independent convolutions per launch.
 *
 * No counterpart in the reference, which runs one convolution per call
 * (src/convGemm.h:20-64).  A batch groups its problems by the kernel
 * configuration their plans use (tuning database, cost model or analytic
 * planner, as for convgemm_conv_ex) and runs each group as ONE grouped
 * launch whose workgroups walk the tiles of all the group's problems, so that
 * small and uneven layers fill the GPU together.  With
 * CONVGEMM_BATCH_GRAPH the launches are captured once into a hipGraph at
 * creation and every run replays it.  A run is one element-wise launch
 * (weight repacks, beta-only scalings), the grouped implicit GEMMs and one
 * split-K reduction.  With CONVGEMM_BATCH_SINGLE the forward NHWC problems
 * run in ONE persistent launch whatever their tile shapes; with static
 * weights and a graph, a batch of such problems is one kernel per run.
 */
#ifndef CONVGEMM_AMD_CONVGEMM_BATCH_H
#define CONVGEMM_AMD_CONVGEMM_BATCH_H

#include "convgemm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One problem: the arguments of convgemm_conv_ex.  All problems of a batch
 * share one operand type; their outputs must not overlap each other or any
 * operand (problems run concurrently). */
typedef struct {
    convgemm_desc_t desc;
    float alpha, beta;
    const void *in;
    const void *x;
    void *out;
    convgemm_epilogue_t epi;  /* forward only; NULL members are absent */
    float scales[2];          /* FP8: {scale_in, scale_x}; ignored otherwise */
} convgemm_problem_t;

typedef struct convgemm_batch_s *convgemm_batch_t;

enum {
    /* capture the run's launches into a hipGraph at creation; runs replay it */
    CONVGEMM_BATCH_GRAPH = 1,
    /* persistent workgroups: each group launches as many workgroups as are
     * resident at once, and each walks the group's tiles with a fixed stride */
    CONVGEMM_BATCH_PERSISTENT = 2,
    /* repack the weights once, at creation: later changes to the weights of
     * problems that need a repack (forward except NCHW 1x1, NCHW backward
     * data) are not seen by runs */
    CONVGEMM_BATCH_STATIC_WEIGHTS = 4,
    /* at most 16 grouped implicit-GEMM launches (18 kernels per run): when the
     * plans name more kernel configurations, the problems of the groups with
     * the least work run on the configuration of a larger group of their
     * operation.  Those problems' outputs then differ from a per-call run in
     * summation order (within the operand type's accumulation tolerance);
     * without this flag every output is bit-identical to a per-call run. */
    CONVGEMM_BATCH_BOUNDED = 8,
    /* one persistent launch for every forward NHWC problem whose channels
     * come in 16-byte chunks (FP32: C % 4 == 0, BF16: C % 8, FP8: C % 16):
     * each is planned on one of the tile shapes of the single-launch kernel
     * (4-wave pipelined tiles of 64x64 to 256x64, any raster order, no
     * split-K) and the kernel's workgroups walk the tiles of all of them,
     * longest K first.  Other problems are grouped as without the flag.
     * Outputs are bit-identical to per-call runs of the same plans
     * (convgemm_batch_plan). */
    CONVGEMM_BATCH_SINGLE = 16
};

/* Plan the n problems, allocate the batch's device workspace (repacked
 * weights, split-K partials, the problem table) and, with
 * CONVGEMM_BATCH_GRAPH, capture the graph.  The operand pointers are bound
 * at creation: runs read and write the same buffers.  A problem whose batch
 * needs more than one 2 GiB addressing window is rejected
 * (CONVGEMM_ERR_UNSUPPORTED; run it with convgemm_conv_ex).  Problems that
 * are no-ops under the quick-return rules of convgemm_conv_ex are dropped. */
convgemm_status_t convgemm_batch_create(const convgemm_problem_t *problems, int n, convgemm_dtype_t dtype,
                                        unsigned flags, convgemm_batch_t *batch);

/* Run every problem of the batch, asynchronously on `stream`.  Runs of one
 * batch share its workspace: runs that overlap in time (on different
 * streams) race. */
convgemm_status_t convgemm_batch_run(convgemm_batch_t batch, convgemm_stream_t stream);

/* Launch structure: grouped implicit-GEMM launches, kernels per run (all
 * kinds), and whether runs replay a graph.  Any pointer may be NULL. */
convgemm_status_t convgemm_batch_info(convgemm_batch_t batch, int *groups, int *kernels_per_run, int *graph);

/* The plan problem i (its index in the list given at creation) runs: tile
 * configuration and split-K factor as convgemm_plan reports them; cfg -1 for
 * a problem with no implicit GEMM (a no-op, or alpha == 0). */
convgemm_status_t convgemm_batch_plan(convgemm_batch_t batch, int i, int *cfg, int *splits);

/* Waits for the device, then frees the batch. NULL is ignored. */
void convgemm_batch_destroy(convgemm_batch_t batch);

#ifdef __cplusplus
}
#endif

#endif /* CONVGEMM_AMD_CONVGEMM_BATCH_H */
