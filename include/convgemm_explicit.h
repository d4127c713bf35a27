/*
 * convgemm_explicit.h -- the explicit path around the implicit GEMM:
 * patch-matrix transforms, a plain GEMM on the same MFMA kernels, and layout
 * converters.  All pointers are device pointers.
 */
#ifndef CONVGEMM_AMD_CONVGEMM_EXPLICIT_H
#define CONVGEMM_AMD_CONVGEMM_EXPLICIT_H

#include "convgemm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Explicit transforms: same names, argument order and semantics as the      */
/* reference's, on device pointers; like the drop-in convolutions they run   */
/* on the library stream and return when it has finished.  Errors are        */
/* recorded for convgemm_get_last_error().                                    */
/* ------------------------------------------------------------------------ */

/* Replaces im2row_nhwc (src/im2row_nhwc.h, src/im2row_nhwc.c:147-204):
 * rows[p*ld + r] = in[b][ix][iy][c] for every in-bounds tap, p =
 * (b*oheight + x)*owidth + y, r = (c*kheight + kx)*kwidth + ky; padded taps
 * are left untouched (the caller zero-fills).  The reference's active branch
 * ignores ld and uses c*kheight*kwidth; every reference caller passes that. */
void im2row_nhwc(float *rows, int ld, const float *in, int batches, int height, int width, int channels,
                 int oheight, int owidth, int kheight, int kwidth, int vpadding, int hpadding, int vstride,
                 int hstride, int vdilation, int hdilation);

/* Replaces row2im_nhwc (src/im2row_nhwc.c:206-229): out[b][ix][iy][c] +=
 * rows[p*ld + r] over every (p, r) that maps to it; m and n are unused, as
 * in the reference. */
void row2im_nhwc(int m, int n, const float *rows, int ld, float *out, int batches, int height, int width,
                 int channels, int oheight, int owidth, int kheight, int kwidth, int vpadding, int hpadding,
                 int vstride, int hstride, int vdilation, int hdilation);

/* Replaces im2col_nchw (src/im2col_nchw.h:20, src/im2col_nchw.c:146-204):
 * cols[r*ld + p] = in[b][c][ix][iy] for in-bounds taps (reference callers
 * pass ld = batches*oheight*owidth). */
void im2col_nchw(float *cols, int ld, const float *in, int batches, int channels, int height, int width,
                 int oheight, int owidth, int kheight, int kwidth, int vpadding, int hpadding, int vstride,
                 int hstride, int vdilation, int hdilation);

/* Replaces col2im_nchw (src/im2col_nchw.c:337-359): out[b][c][ix][iy] +=
 * cols[r*ld + p]. */
void col2im_nchw(int m, int n, const float *cols, int ld, float *out, int batches, int channels, int height,
                 int width, int oheight, int owidth, int kheight, int kwidth, int vpadding, int hpadding,
                 int vstride, int hstride, int vdilation, int hdilation);

/* Replaces transpose_nchw (src/im2col_nchw.c:206-233):
 * out[((b*kn + start_col + j)*ho + x)*wo + y] = in[j*ld + i] (beta == 0) or
 * beta*out[...] + in[j*ld + i], (b, x, y) = pixel start_row + i, i < rows,
 * j < cols. */
void transpose_nchw(int rows, int cols, const float *in, int ld, float beta, float *out, int kn, int ho, int wo,
                    int start_row, int start_col);

/* Extended form of im2row_nhwc / im2col_nchw (no counterpart in the
 * reference, whose transforms are FP32 only): the patch matrix of a
 * descriptor's input with FP32 or BF16 elements (dtype; the copy does no
 * arithmetic), asynchronous on `stream`.  NHWC: rows[p*ld + r] with ld >=
 * c*kh*kw; NCHW: cols[r*ld + p] with ld >= b*ho*wo; p and r as above.  Only
 * in-bounds taps are written (the caller zero-fills).  The weight gradient's
 * and backward data's descriptors describe the same input. */
convgemm_status_t convgemm_im2row_ex(const convgemm_desc_t *desc, convgemm_dtype_t dtype, const void *x, void *out,
                                     long long ld, convgemm_stream_t stream);

/* Extended form of transpose_nchw: the same operation, asynchronous on
 * `stream` (the drop-in form waits for the library stream). */
convgemm_status_t convgemm_transpose_nchw_ex(int rows, int cols, const float *in, long long ld, float beta,
                                            float *out, int kn, int ho, int wo, int start_row, int start_col,
                                            convgemm_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Plain GEMM and layout converters (asynchronous on `stream`)                */
/* ------------------------------------------------------------------------ */

/* C = alpha*op(A)*op(B) + beta*C with op(A) m x k, op(B) k x n, C m x n:
 * the reference's gemm_blis_B3A2C0 (src/gemm_blis.h:90-94,
 * src/gemm_blis_B3A2C0.c:28-117) used as a plain GEMM (generic packing, no
 * post-process, no convolution), run on the implicit-GEMM kernels as a 1x1
 * convolution.  order*: 'C' column-major or 'R' row-major; trans*: 'N' or
 * 'T'.  Operands whose layout is not the kernel's are packed into an
 * internal workspace first.  dtype: operand type of A and B (C is FP32);
 * scales: FP8 {scale_A, scale_B}, else NULL.  Quick-return rules as the
 * reference: nothing when m == 0, n == 0, or (alpha == 0 or k == 0) and
 * beta == 1; C = beta*C when alpha == 0. */
convgemm_status_t convgemm_gemm(char orderA, char orderB, char orderC, char transA, char transB, int m, int n, int k,
                                float alpha, const void *A, int ldA, const void *B, int ldB, float beta, float *C,
                                int ldC, convgemm_dtype_t dtype, const float *scales, convgemm_stream_t stream);

/* Activations [b][c][h][w] <-> [b][h][w][c]; esize 4, 2 or 1 bytes. */
convgemm_status_t convgemm_nchw_to_nhwc(const void *src, void *dst, int b, int c, int h, int w, int esize,
                                        convgemm_stream_t stream);
convgemm_status_t convgemm_nhwc_to_nchw(const void *src, void *dst, int b, int h, int w, int c, int esize,
                                        convgemm_stream_t stream);
/* Weights: NHWC-family [c][kh][kw][kn] <-> NCHW-family [kn][c][kh][kw]. */
convgemm_status_t convgemm_weights_nhwc_to_nchw(const void *src, void *dst, int c, int kh, int kw, int kn,
                                                int esize, convgemm_stream_t stream);
convgemm_status_t convgemm_weights_nchw_to_nhwc(const void *src, void *dst, int kn, int c, int kh, int kw,
                                                int esize, convgemm_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CONVGEMM_AMD_CONVGEMM_EXPLICIT_H */
