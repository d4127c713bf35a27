/*
 * lowp.c -- BF16 / FP8-e4m3 operand paths of convgemm_conv_ex.
 */
#include "cg_host.h"

convgemm_status_t cg_run_lowp(const convgemm_desc_t *d, float alpha, float beta, const void *in, const void *x,
                              void *out, const convgemm_epilogue_t *epi, convgemm_dtype_t dtype,
                              const float *scales, void *stream) {
    (void)d; (void)alpha; (void)beta; (void)in; (void)x; (void)out; (void)epi; (void)scales; (void)stream;
    (void)dtype;
    return cg_set_error(CONVGEMM_ERR_UNSUPPORTED, "operand type not supported yet");
}

size_t cg_lowp_extra_bytes(const convgemm_desc_t *d, convgemm_dtype_t dtype) {
    (void)d; (void)dtype;
    return 0;
}
