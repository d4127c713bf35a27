/*
 * cg_host.h -- internal interfaces of the C host layer (not installed).
 */
#ifndef CONVGEMM_AMD_CG_HOST_H
#define CONVGEMM_AMD_CG_HOST_H

#include <stddef.h>

#include "convgemm.h"
#include "../cg_internal.h"

/* Size of each buffer handed out by alloc_pack_buffs. */
#define CG_PACK_BYTES (64 << 20)

/* Planner hook: given the operand type and the problem (tile fields unset),
 * choose a tile configuration and a split-K factor (cfg < 0: use the default). */
typedef void (*cg_plan_hook_t)(void *ctx, int elem, const cg_kparams *kp, int *cfg, int *splits);

void cg_set_plan_hook(cg_plan_hook_t hook, void *ctx);
/* Force every call to use (cfg, splits); cfg < 0 clears it. Tuning hook. */
void cg_force_config(int cfg, int splits);
void cg_fill_tiles(cg_kparams *kp, const cg_tilecfg *t, int splits);
/* Does the problem's largest halo fit a halo-tile configuration (1 for the others)? */
int cg_halo_fits(const cg_kparams *kp, const cg_tilecfg *t);
/* Configuration cfg serves the problem (operation, loads, halo). */
int cg_plan_usable(int elem, int cfg, const cg_kparams *kp);
/* The analytic ("no learning") planner. */
void cg_default_plan(int elem, const cg_kparams *base, int *cfg, int *splits);
/* Times of given candidates from the planner's evidence (tuning database
 * timings, else the cost model; 1 if us[] was filled). */
int cg_plan_times(int elem, const cg_kparams *kp, const int *cfgs, const int *splits, int n, float *us);
/* The plan forced through cg_force_config (1), or none (0). */
int cg_forced_config(int *cfg, int *splits);
size_t cg_partial_bytes(const cg_kparams *kp);
/* Plan chosen by the last call on this thread: tile configuration, split-K
 * factor and number of kernels launched. */
void cg_last_plan(int *cfg, int *splits, int *nkernels);
/* Test hook: the next n device allocations made by the library fail. */
void cg_debug_fail_next_allocs(int n);

convgemm_status_t cg_set_error(convgemm_status_t st, const char *msg);

/* Validate a call and describe it as kernel parameters (tile fields unset,
 * B/P of the first batch chunk), the element type and the batch chunk.
 * wsw: the weight workspace the call will use (NULL: internal, aligned). */
convgemm_status_t cg_prepare(const convgemm_desc_t *d, convgemm_dtype_t dtype, const void *in, const void *x,
                             const void *wsw, cg_kparams *kp, int *elem, int *chunk);
/* Candidate plans of a prepared problem: every supported configuration with
 * split-K factors that leave >= 2 K tiles per split, plus one stream-K plan
 * (splits 0) per configuration.  At most (split factors + 1) x configurations,
 * which CONVGEMM_MAX_CANDIDATES bounds.  Returns the count. */
int cg_candidates(int elem, const cg_kparams *kp, int *cfgs, int *splits, int max);

/* Plan (cfg, splits) as a call would: forced plan, planner hook, analytic model. */
int cg_choose_plan(int elem, const cg_kparams *kp, int *cfg, int *splits);
/* Weight repack the kernels need for this call (arguments of
 * cg_launch_permute3); 0 when none. */
int cg_weight_repack(const convgemm_desc_t *d, int elem, int *d0, int *d1, int *d2, int *perm);
int cg_esize(int elem);
/* Device allocation through the library's allocator (honours the test hook). */
int cg_malloc(void **p, size_t bytes);
convgemm_status_t cg_hip_fail(int err, const char *what);
/* Stream-K: workgroups for a problem with tile fields filled (0: none for
 * this configuration/operation, or no tile left to cut), and the bytes of
 * their partial tiles.  cg_sk_split also gives the data-parallel tiles that
 * precede the stream-K region. */
int cg_sk_groups(int elem, int cfg, const cg_kparams *kp);
int cg_sk_split(int elem, int cfg, const cg_kparams *kp, int *dp);
size_t cg_sk_bytes(int groups, const cg_tilecfg *t);
/* The library stream of the drop-in calls (hipStream_t). */
void *cg_get_stream(void);

/* Backward data of NHWC convolutions with unit strides (and padding within
 * the kernel's reach) runs as a forward convolution of dY with the kernel
 * turned by 180 degrees: cg_forward_form rewrites such a prepared problem
 * (Ho, Wo: dY's extent) in place and sets kp->ffd; cg_dgrad_forward_form_ok
 * says whether a prepared backward-data problem takes that form. */
void cg_forward_form(cg_kparams *kp);
int cg_dgrad_forward_form_ok(const cg_kparams *kp);

/* Is `stream` (hipStream_t) being captured into a graph? */
int cg_stream_capturing(void *stream);

/* A grow-only internal device buffer (see cg_growbuf_get in convgemm.c). */
typedef struct {
    void *ptr;
    size_t bytes;
    int retired;      /* outgrown buffers kept allocated */
    const char *what; /* name used in error messages */
} cg_growbuf;
/* At least `need` bytes of gb for work queued on `stream`; never frees a
 * buffer, and fails with CONVGEMM_ERR_WORKSPACE instead of growing while the
 * stream is being captured. */
convgemm_status_t cg_growbuf_get(cg_growbuf *gb, size_t need, void *stream, void **out);

/* Test hook: bytes one batch chunk may address (0 restores 2 GiB).  Smaller
 * values make small problems run the batch-chunking path. */
void cg_set_chunk_window(long long bytes);

/* Test hook: run backward data of unit-stride NHWC problems in its forward
 * form (1, the default) or on the residue-class kernel (0). */
void cg_set_dgrad_forward(int on);

#endif
