/*
 * costmodel.c -- learned cost model (placeholder until the MLP lands).
 */
#include "convgemm_tune.h"
#include "../cg_internal.h"
#include "cg_host.h"

int cg_costmodel_choose(int elem, const cg_kparams *kp, int *cfg, int *splits) {
    (void)elem; (void)kp; (void)cfg; (void)splits;
    return -1;
}

convgemm_status_t convgemm_costmodel_train(int epochs, float lr, unsigned seed, float *loss, convgemm_stream_t stream) {
    (void)epochs; (void)lr; (void)seed; (void)loss; (void)stream;
    return cg_set_error(CONVGEMM_ERR_UNSUPPORTED, "cost model not built");
}
convgemm_status_t convgemm_costmodel_save(const char *path) { (void)path; return cg_set_error(CONVGEMM_ERR_UNSUPPORTED, "cost model not built"); }
convgemm_status_t convgemm_costmodel_load(const char *path) { (void)path; return cg_set_error(CONVGEMM_ERR_UNSUPPORTED, "cost model not built"); }
convgemm_status_t convgemm_costmodel_predict(const convgemm_desc_t *desc, convgemm_dtype_t dtype, int *cfgs,
                                             int *splits, float *us, int *n_out) {
    (void)desc; (void)dtype; (void)cfgs; (void)splits; (void)us; (void)n_out;
    return cg_set_error(CONVGEMM_ERR_UNSUPPORTED, "cost model not built");
}
