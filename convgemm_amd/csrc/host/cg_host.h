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

/* Planner hook: given the problem (tile fields unset), choose a tile
 * configuration and a split-K factor (cfg < 0: fall back to the default). */
typedef void (*cg_plan_hook_t)(void *ctx, const cg_kparams *kp, int *cfg, int *splits);

void cg_set_plan_hook(cg_plan_hook_t hook, void *ctx);
/* Force every call to use (cfg, splits); cfg < 0 clears it. Tuning hook. */
void cg_force_config(int cfg, int splits);
void cg_fill_tiles(cg_kparams *kp, const cg_tilecfg *t, int splits);
size_t cg_partial_bytes(const cg_kparams *kp);
/* Plan chosen by the last call on this thread: tile configuration, split-K
 * factor and number of kernels launched. */
void cg_last_plan(int *cfg, int *splits, int *nkernels);
/* Test hook: the next n device allocations made by the library fail. */
void cg_debug_fail_next_allocs(int n);

convgemm_status_t cg_set_error(convgemm_status_t st, const char *msg);

/* BF16 / FP8 operand paths (lowp.c). */
convgemm_status_t cg_run_lowp(const convgemm_desc_t *d, float alpha, float beta, const void *in, const void *x,
                              void *out, const convgemm_epilogue_t *epi, convgemm_dtype_t dtype,
                              const float *scales, void *stream);
size_t cg_lowp_extra_bytes(const convgemm_desc_t *d, convgemm_dtype_t dtype);

#endif
