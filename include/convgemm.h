/*
 * convgemm.h -- MI355X (gfx950) convolution as implicit GEMM.
 *
 * Drop-in for convGemm's C interface (src/convGemm.h:20-64): the same five
 * symbols with the same argument order, layouts and semantics, except that
 * every tensor pointer is a DEVICE pointer (hipMalloc'd memory on the current
 * HIP device).  Each drop-in call launches on the library stream (the null
 * stream unless convgemm_set_stream() chose another) and returns after the
 * stream has finished, matching the reference's blocking behaviour.
 *
 * Layouts (identical to the reference):
 *   NHWC: X, dX [b][h][w][c]   W, dW [c][kh][kw][kn]   Y, dY [b][ho][wo][kn]
 *   NCHW: X, dX [b][c][h][w]   W, dW [kn][c][kh][kw]   Y, dY [b][kn][ho][wo]
 *   ho = (h + 2*vpadding - vdilation*(kh-1) - 1)/vstride + 1, likewise wo.
 *
 * Error reporting: the drop-in functions return void like the reference;
 * a rejected call (bad shape, NULL operand, HIP failure) writes nothing and
 * records a status readable with convgemm_get_last_error() and
 * convgemm_last_error_message().
 */
#ifndef CONVGEMM_AMD_CONVGEMM_H
#define CONVGEMM_AMD_CONVGEMM_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Drop-in entry points (replace src/convGemm.h:20-64)                       */
/* ------------------------------------------------------------------------ */

/* Replaces alloc_pack_buffs (src/convGemm.h:20, src/convGemm.c:30-40).
 * Allocates the two device workspaces used by the conv calls (split-K
 * partial sums and the repacked weights).  Returns 0 on success, 1 if an
 * allocation failed.  Release them with convgemm_free_pack_buffs().
 * The conv calls use ac_pack / bc_pack only when they came from this
 * function (the library knows their size); any other pointer, like NULL,
 * makes the call use the library's internal workspace. */
int alloc_pack_buffs(float **Ac_pack, float **Bc_pack);

/* Replaces sconvGemmNHWC (src/convGemm.h:22-33, src/convGemm.c:48-77).
 * trans == 'N': out = epi(conv(x, in)): in = W, x = X, out = Y, with the
 *   optional fused epilogue t += bias; t = (t-mean)*inv_std*gamma + beta;
 *   t = max(t, 0).  A NULL epilogue pointer means the term is absent.
 * otherwise:    out = dW = im2row(x)^T * in: in = dY, x = X (epilogue
 *   arguments ignored). */
void sconvGemmNHWC(char trans,
                   int b, int h, int w, int c,
                   int kn, int kh, int kw,
                   int vpadding, int hpadding,
                   int vstride, int hstride,
                   int vdilation, int hdilation,
                   const float *in,
                   const float *x,
                   float *out, const float *bias_vector,
                   const float *bn_running_mean, const float *bn_inv_std,
                   const float *bn_gamma, const float *bn_beta, bool relu,
                   float *ac_pack, float *bc_pack);

/* Replaces sconvGemmNHWC_back (src/convGemm.h:35-42, src/convGemm.c:82-97).
 * dx += row2im(weights^T * dy).  NOTE: stride comes before padding here. */
void sconvGemmNHWC_back(int b, int h, int w, int c,
                        int kn, int kh, int kw,
                        int vstride, int hstride,
                        int vpadding, int hpadding,
                        int vdilation, int hdilation,
                        const float *weights,
                        const float *dy, float *dx,
                        float *ac_pack, float *bc_pack);

/* Replaces sconvGemmNCHW (src/convGemm.h:44-55, src/convGemm.c:99-125). */
void sconvGemmNCHW(char trans,
                   int b, int c, int h, int w,
                   int kn, int kh, int kw,
                   int vpadding, int hpadding,
                   int vstride, int hstride,
                   int vdilation, int hdilation,
                   const float *in,
                   const float *x,
                   float *out, const float *bias_vector,
                   const float *bn_running_mean, const float *bn_inv_std,
                   const float *bn_gamma, const float *bn_beta, bool relu,
                   float *ac_pack, float *bc_pack);

/* Replaces sconvGemmNCHW_back (src/convGemm.h:57-64, src/convGemm.c:130-146).
 * dx += col2im(dy * weights) (accumulates, like the reference). */
void sconvGemmNCHW_back(int b, int c, int h, int w,
                        int kn, int kh, int kw,
                        int vstride, int hstride,
                        int vpadding, int hpadding,
                        int vdilation, int hdilation,
                        const float *weights,
                        const float *dy, float *dx,
                        float *ac_pack, float *bc_pack);

/* ------------------------------------------------------------------------ */
/* Extended API                                                               */
/* ------------------------------------------------------------------------ */

typedef enum {
    CONVGEMM_SUCCESS = 0,
    CONVGEMM_ERR_INVALID_VALUE = 1,  /* bad shape, stride, trans, NULL operand */
    CONVGEMM_ERR_ALLOC = 2,          /* device allocation failed */
    CONVGEMM_ERR_HIP = 3,            /* a HIP runtime call failed */
    CONVGEMM_ERR_WORKSPACE = 4,      /* workspace too small for the call */
    CONVGEMM_ERR_UNSUPPORTED = 5     /* valid but not supported (e.g. kernel > 64 taps per side) */
} convgemm_status_t;

typedef enum { CONVGEMM_FWD = 0, CONVGEMM_WGRAD = 1, CONVGEMM_DGRAD = 2 } convgemm_op_t;
typedef enum { CONVGEMM_NHWC = 0, CONVGEMM_NCHW = 1 } convgemm_layout_t;

/* Operand storage type of `in` and `x`.  Accumulation is always FP32 and
 * `out` is always FP32.  CONVGEMM_BF16: operands are bfloat16 (raw uint16).
 * CONVGEMM_FP8_E4M3: operands are OCP e4m3fn bytes holding value/scale with
 * per-tensor scales given in convgemm_conv_ex's `scales`. */
typedef enum { CONVGEMM_F32 = 0, CONVGEMM_BF16 = 1, CONVGEMM_FP8_E4M3 = 2 } convgemm_dtype_t;

typedef struct {
    int op;      /* convgemm_op_t */
    int layout;  /* convgemm_layout_t */
    int b, h, w, c, kn, kh, kw;
    int vpad, hpad, vstride, hstride, vdil, hdil;
} convgemm_desc_t;

typedef struct {
    const float *bias, *mean, *inv_std, *gamma, *beta;
    int relu;
} convgemm_epilogue_t;

/* hipStream_t, kept opaque so C callers need no HIP header. */
typedef void *convgemm_stream_t;

/* General form of the six operations, asynchronous on `stream`:
 *   FWD:   out(Y)  = epi(alpha*conv(x=X, in=W) + beta*Y)
 *   WGRAD: out(dW) = alpha*im2row(x=X)^T in(dY) + beta*dW
 *   DGRAD: out(dX) = alpha*row2im(in(W)^T x(dY)) + beta*dX
 * (the drop-in calls are alpha=1, beta=0 for FWD/WGRAD and beta=1 for DGRAD).
 * alpha == 0 follows quick_gemm.h: out = beta*out (0 if beta == 0) and the
 * epilogue is not applied.  `epi` may be NULL.  `workspace`/`ws_bytes` may
 * be NULL/0, in which case the library uses an internal device buffer: one
 * grow-only buffer shared by every call that passes no workspace, so such
 * calls on different streams must not overlap in time.  An outgrown internal
 * buffer is kept allocated (graphs captured earlier may still use it), and
 * while `stream` is being captured the buffer does not grow: the call then
 * returns CONVGEMM_ERR_WORKSPACE, and calls made under capture pass a
 * workspace of convgemm_workspace_size() bytes.
 * `scales` (FP8 only): {scale_in, scale_x}, the true values are
 * stored*scale. */
convgemm_status_t convgemm_conv_ex(const convgemm_desc_t *desc, float alpha, float beta,
                                   const void *in, const void *x, void *out,
                                   const convgemm_epilogue_t *epi, convgemm_dtype_t dtype,
                                   const float *scales,
                                   void *workspace, size_t ws_bytes, convgemm_stream_t stream);

/* No counterpart in the reference: convgemm_conv_ex for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc).  It makes no device allocation and no
 * free: its repacked weights and split-K partials live in a static device
 * array of convgemm_noalloc_scratch_bytes() (32 MiB) per GPU, held until the
 * process ends.  Weights whose repacked copy exceeds the scratch are rejected
 * with CONVGEMM_ERR_WORKSPACE; plans whose partials do not fit use fewer
 * split-K slices.  Plans come from a forced plan, the tuning database or the
 * cost model's cached choices, else the analytic planner (no model inference,
 * which allocates).  Every call on one GPU shares that GPU's scratch, so calls
 * that overlap in time on one GPU (other streams, other threads) race. */
convgemm_status_t convgemm_conv_noalloc(const convgemm_desc_t *desc, float alpha, float beta,
                                        const void *in, const void *x, void *out,
                                        const convgemm_epilogue_t *epi, convgemm_dtype_t dtype,
                                        const float *scales, convgemm_stream_t stream);
/* Bytes of the static device scratch behind convgemm_conv_noalloc. */
size_t convgemm_noalloc_scratch_bytes(void);

/* Weight prepacking for repeated calls with the same weights (no counterpart
 * in the reference, which packs on every call).  Some operations read the
 * weights in another order than the caller's (forward: [kn][kh][kw][c],
 * except NCHW 1x1 kernels, already in that order; NCHW backward data:
 * [c][kh][kw][kn]); every call of
 * convgemm_conv_ex then starts with a repack kernel.  convgemm_pack_weights
 * does that repack once into `packed` (convgemm_packed_weights_size bytes;
 * 0 means the operation reads the weights as given), and convgemm_conv_packed
 * is convgemm_conv_ex on such packed weights (`workspace` then only holds
 * split-K partials).  Packed weights stay valid while the weights and the
 * descriptor's shape are unchanged. */
size_t convgemm_packed_weights_size(const convgemm_desc_t *desc, convgemm_dtype_t dtype);
convgemm_status_t convgemm_pack_weights(const convgemm_desc_t *desc, convgemm_dtype_t dtype, const void *in,
                                        void *packed, convgemm_stream_t stream);
convgemm_status_t convgemm_conv_packed(const convgemm_desc_t *desc, float alpha, float beta,
                                       const void *packed, const void *x, void *out,
                                       const convgemm_epilogue_t *epi, convgemm_dtype_t dtype,
                                       const float *scales, void *workspace, size_t ws_bytes,
                                       convgemm_stream_t stream);

/* Convert n FP32 values on the device to `dtype`, asynchronously on `stream`.
 * CONVGEMM_BF16: round to nearest even, *scale_out = 1.
 * CONVGEMM_FP8_E4M3: per-tensor amax scaling, scale = amax/448 (1 if amax is
 *   0), stored = e4m3fn(x/scale) rounded to nearest even; *scale_out receives
 *   the scale (the call then waits for the stream).  scale_out may be NULL. */
convgemm_status_t convgemm_quantize(const float *src, void *dst, size_t n, convgemm_dtype_t dtype,
                                    float *scale_out, convgemm_stream_t stream);

/* Device workspace bytes convgemm_conv_ex needs for this problem. */
size_t convgemm_workspace_size(const convgemm_desc_t *desc, convgemm_dtype_t dtype);

/* Output spatial size with the reference's formula (src/convGemm.c:60-61). */
int convgemm_out_size(int in, int pad, int dil, int k, int stride);

/* Frees buffers from alloc_pack_buffs (the reference has no free function;
 * its buffers came from aligned_alloc). */
void convgemm_free_pack_buffs(float *Ac_pack, float *Bc_pack);

/* Stream used by the drop-in calls (default: the null stream). */
convgemm_status_t convgemm_set_stream(convgemm_stream_t stream);

convgemm_status_t convgemm_get_last_error(void);
const char *convgemm_last_error_message(void);
const char *convgemm_status_string(convgemm_status_t status);

/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int convgemm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CONVGEMM_AMD_CONVGEMM_H */
