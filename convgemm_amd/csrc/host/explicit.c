/*
 * explicit.c -- explicit transforms, plain GEMM and layout converters
 * (include/convgemm_explicit.h).
 *
 * The GEMM runs on the implicit-GEMM kernels as a 1x1 convolution in NHWC:
 * Y[P][N'] = X[P][K] W[K][N'] with every operand row-major and packed.  A
 * row-major C takes P = m, N' = n, X = op(A), W = op(B); a column-major C is
 * the row-major C^T = op(B)^T op(A)^T, so P = n, N' = m, X = op(B)^T,
 * W = op(A)^T.  An operand whose storage is not that packed row-major
 * matrix is copied (strided) or transposed into an internal workspace.
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "convgemm.h"
#include "convgemm_explicit.h"
#include "../cg_internal.h"
#include "cg_host.h"

#define AL256(x) (((size_t)(x) + 255) & ~(size_t)255)

/* Grow-only packing workspace of the GEMM entry (cg_growbuf_get: outgrown
 * buffers are retired, not freed, and it does not grow under capture).
 * Concurrent GEMM calls that pack share it and must be ordered by the caller
 * (one stream), like the reference's shared pack buffers. */
static cg_growbuf e_ws = {NULL, 0, 0, "GEMM packing workspace"};

/* ------------------------------------------------------------- transforms */

static int tshape(cg_tshape *s, int b, int h, int w, int c, int ho, int wo, int kh, int kw, int vp, int hp, int vs,
                  int hs, int vd, int hd) {
    if (b < 0 || h < 0 || w < 0 || c < 0 || ho < 0 || wo < 0 || kh < 1 || kw < 1 || vs < 1 || hs < 1 || vd < 1 ||
        hd < 1) {
        cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "invalid transform shape");
        return 0;
    }
    s->b = b; s->h = h; s->w = w; s->c = c; s->ho = ho; s->wo = wo; s->kh = kh; s->kw = kw;
    s->vp = vp; s->hp = hp; s->vs = vs; s->hs = hs; s->vd = vd; s->hd = hd;
    return 1;
}

static void finish(int e, const char *what) {
    if (e) {
        cg_hip_fail(e, what);
        return;
    }
    hipError_t r = hipStreamSynchronize((hipStream_t)cg_get_stream());
    if (r != hipSuccess) cg_hip_fail(r, what);
    else cg_set_error(CONVGEMM_SUCCESS, "");
}

void im2row_nhwc(float *rows, int ld, const float *in, int batches, int height, int width, int channels,
                 int oheight, int owidth, int kheight, int kwidth, int vpadding, int hpadding, int vstride,
                 int hstride, int vdilation, int hdilation) {
    cg_tshape s;
    if (!tshape(&s, batches, height, width, channels, oheight, owidth, kheight, kwidth, vpadding, hpadding, vstride,
                hstride, vdilation, hdilation))
        return;
    if (ld < channels * kheight * kwidth) {
        cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "im2row_nhwc: ld < c*kh*kw");
        return;
    }
    finish(cg_launch_im2row(&s, in, rows, ld, 0, 4, cg_get_stream()), "im2row_nhwc");
}

void row2im_nhwc(int m, int n, const float *rows, int ld, float *out, int batches, int height, int width,
                 int channels, int oheight, int owidth, int kheight, int kwidth, int vpadding, int hpadding,
                 int vstride, int hstride, int vdilation, int hdilation) {
    (void)m;
    (void)n;
    cg_tshape s;
    if (!tshape(&s, batches, height, width, channels, oheight, owidth, kheight, kwidth, vpadding, hpadding, vstride,
                hstride, vdilation, hdilation))
        return;
    if (ld < channels * kheight * kwidth) {
        cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "row2im_nhwc: ld < c*kh*kw");
        return;
    }
    finish(cg_launch_row2im(&s, rows, ld, out, 0, cg_get_stream()), "row2im_nhwc");
}

void im2col_nchw(float *cols, int ld, const float *in, int batches, int channels, int height, int width,
                 int oheight, int owidth, int kheight, int kwidth, int vpadding, int hpadding, int vstride,
                 int hstride, int vdilation, int hdilation) {
    cg_tshape s;
    if (!tshape(&s, batches, height, width, channels, oheight, owidth, kheight, kwidth, vpadding, hpadding, vstride,
                hstride, vdilation, hdilation))
        return;
    if ((long long)ld < (long long)batches * oheight * owidth) {
        cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "im2col_nchw: ld < b*ho*wo");
        return;
    }
    finish(cg_launch_im2row(&s, in, cols, ld, 1, 4, cg_get_stream()), "im2col_nchw");
}

void col2im_nchw(int m, int n, const float *cols, int ld, float *out, int batches, int channels, int height,
                 int width, int oheight, int owidth, int kheight, int kwidth, int vpadding, int hpadding,
                 int vstride, int hstride, int vdilation, int hdilation) {
    (void)m;
    (void)n;
    cg_tshape s;
    if (!tshape(&s, batches, height, width, channels, oheight, owidth, kheight, kwidth, vpadding, hpadding, vstride,
                hstride, vdilation, hdilation))
        return;
    if ((long long)ld < (long long)batches * oheight * owidth) {
        cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "col2im_nchw: ld < b*ho*wo");
        return;
    }
    finish(cg_launch_row2im(&s, cols, ld, out, 1, cg_get_stream()), "col2im_nchw");
}

void transpose_nchw(int rows, int cols, const float *in, int ld, float beta, float *out, int kn, int ho, int wo,
                    int start_row, int start_col) {
    if (rows < 0 || cols < 0 || ho < 1 || wo < 1 || start_row < 0 || start_col < 0 || (cols > 0 && ld < rows)) {
        cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "transpose_nchw: invalid arguments");
        return;
    }
    finish(cg_launch_transpose_nchw(rows, cols, in, ld, beta, out, kn, ho, wo, start_row, start_col,
                                    cg_get_stream()),
           "transpose_nchw");
}

convgemm_status_t convgemm_transpose_nchw_ex(int rows, int cols, const float *in, long long ld, float beta,
                                            float *out, int kn, int ho, int wo, int start_row, int start_col,
                                            convgemm_stream_t stream) {
    if (rows < 0 || cols < 0 || ho < 1 || wo < 1 || start_row < 0 || start_col < 0 || (cols > 0 && ld < rows))
        return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "transpose_nchw_ex: invalid arguments");
    int e = cg_launch_transpose_nchw(rows, cols, in, ld, beta, out, kn, ho, wo, start_row, start_col, stream);
    return e ? cg_hip_fail(e, "transpose_nchw_ex") : cg_set_error(CONVGEMM_SUCCESS, "");
}

convgemm_status_t convgemm_im2row_ex(const convgemm_desc_t *d, convgemm_dtype_t dtype, const void *x, void *out,
                                     long long ld, convgemm_stream_t stream) {
    if (!d || (d->layout != CONVGEMM_NHWC && d->layout != CONVGEMM_NCHW))
        return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "im2row_ex: NULL descriptor or unknown layout");
    if (dtype != CONVGEMM_F32 && dtype != CONVGEMM_BF16)
        return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "im2row_ex: FP32 or BF16 elements only");
    const int ho = convgemm_out_size(d->h, d->vpad, d->vdil, d->kh, d->vstride);
    const int wo = convgemm_out_size(d->w, d->hpad, d->hdil, d->kw, d->hstride);
    cg_tshape s;
    if (!tshape(&s, d->b, d->h, d->w, d->c, ho, wo, d->kh, d->kw, d->vpad, d->hpad, d->vstride, d->hstride, d->vdil,
                d->hdil))
        return convgemm_get_last_error();
    const int nchw = d->layout == CONVGEMM_NCHW;
    if (ld < (nchw ? (long long)d->b * ho * wo : (long long)d->c * d->kh * d->kw))
        return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "im2row_ex: ld smaller than a patch-matrix row");
    if ((long long)d->b * ho * wo * d->c * d->kh * d->kw > 0 && (!x || !out))
        return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "im2row_ex: NULL tensor");
    int e = cg_launch_im2row(&s, x, out, ld, nchw, dtype == CONVGEMM_BF16 ? 2 : 4, stream);
    return e ? cg_hip_fail(e, "im2row_ex") : cg_set_error(CONVGEMM_SUCCESS, "");
}

/* ------------------------------------------------------------------- GEMM */

/* op(S) element (i, j) lives at i*ld + j ("row-like") or i + j*ld. */
static int rowlike(char order, char trans) { return (order == 'R') == (trans == 'N'); }

typedef struct {
    const void *p;
    long long ld;
    int rowlike;
    int rows, cols; /* of the matrix the kernel needs */
} opnd_t;

/* Bytes of packing this operand needs (0: usable as stored). */
static size_t pack_need(const opnd_t *o, int es) {
    if (o->rowlike && o->ld == o->cols) return 0;
    return AL256((size_t)o->rows * o->cols * es);
}

static int pack(const opnd_t *o, void *dst, int es, hipStream_t s) {
    if (o->rowlike) return cg_launch_copy2d(o->p, o->ld, dst, o->cols, o->rows, o->cols, es, s);
    /* stored as the row-major transpose (cols x rows, ld) */
    return cg_launch_transpose2d(o->p, o->ld, dst, o->cols, o->cols, o->rows, es, s);
}

convgemm_status_t convgemm_gemm(char orderA, char orderB, char orderC, char transA, char transB, int m, int n, int k,
                                float alpha, const void *A, int ldA, const void *B, int ldB, float beta, float *C,
                                int ldC, convgemm_dtype_t dtype, const float *scales, convgemm_stream_t stream) {
    if ((orderA != 'C' && orderA != 'R') || (orderB != 'C' && orderB != 'R') || (orderC != 'C' && orderC != 'R') ||
        (transA != 'N' && transA != 'T') || (transB != 'N' && transB != 'T'))
        return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "gemm: order must be 'C'/'R' and trans 'N'/'T'");
    if (m < 0 || n < 0 || k < 0) return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "gemm: negative dimension");
    if (dtype != CONVGEMM_F32 && dtype != CONVGEMM_BF16 && dtype != CONVGEMM_FP8_E4M3)
        return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "gemm: unknown dtype");
    const int es = dtype == CONVGEMM_F32 ? 4 : dtype == CONVGEMM_BF16 ? 2 : 1;
    /* quick return (src/gemm_blis_B3A2C0_orig.c:45-47); alpha == 0 and k == 0 follow conv_ex */
    if (m == 0 || n == 0 || ((alpha == 0.0f || k == 0) && beta == 1.0f)) return cg_set_error(CONVGEMM_SUCCESS, "");
    if (!C) return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "gemm: C is NULL");

    /* each stored leading dimension must cover the contiguous extent (BLAS rules) */
    {
        const int a_minld = (orderA == 'C') ? (transA == 'N' ? m : k) : (transA == 'N' ? k : m);
        const int b_minld = (orderB == 'C') ? (transB == 'N' ? k : n) : (transB == 'N' ? n : k);
        const int c_minld = orderC == 'C' ? m : n;
        if (ldA < (a_minld > 1 ? a_minld : 1) || ldB < (b_minld > 1 ? b_minld : 1) ||
            ldC < (c_minld > 1 ? c_minld : 1))
            return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "gemm: leading dimension too small");
    }
    const int ra = rowlike(orderA, transA), rb = rowlike(orderB, transB);
    opnd_t X, W;
    int P, Nn;
    if (orderC == 'R') { /* Y = C: X = op(A) m x k, W = op(B) k x n */
        P = m; Nn = n;
        X = (opnd_t){A, ldA, ra, m, k};
        W = (opnd_t){B, ldB, rb, k, n};
    } else { /* Y = C^T: X = op(B)^T n x k, W = op(A)^T k x m */
        P = n; Nn = m;
        X = (opnd_t){B, ldB, !rb, n, k};
        W = (opnd_t){A, ldA, !ra, k, m};
    }
    const int c_packed = ldC == Nn;
    hipStream_t s = (hipStream_t)stream;
    size_t nx = k > 0 ? pack_need(&X, es) : 0, nw = k > 0 ? pack_need(&W, es) : 0;
    size_t ny = c_packed ? 0 : AL256((size_t)P * Nn * 4);
    void *ws = NULL;
    if (nx + nw + ny) {
        convgemm_status_t st = cg_growbuf_get(&e_ws, nx + nw + ny, s, &ws);
        if (st != CONVGEMM_SUCCESS) return st;
    }
    char *wp = (char *)ws;
    const void *xp = X.p, *wptr = W.p;
    int e;
    if (nx) {
        if ((e = pack(&X, wp, es, s))) return cg_hip_fail(e, "gemm: packing an operand");
        xp = wp;
        wp += nx;
    }
    if (nw) {
        if ((e = pack(&W, wp, es, s))) return cg_hip_fail(e, "gemm: packing an operand");
        wptr = wp;
        wp += nw;
    }
    float *yp = C;
    if (!c_packed) {
        yp = (float *)wp;
        /* C is read when beta != 0, and kept as is when alpha == 0 or k == 0 */
        if ((beta != 0.0f || alpha == 0.0f || k == 0) && (e = cg_launch_copy2d(C, ldC, yp, Nn, P, Nn, 4, s)))
            return cg_hip_fail(e, "gemm: packing C");
    }
    convgemm_desc_t d = {CONVGEMM_FWD, CONVGEMM_NHWC, P, 1, 1, k, Nn, 1, 1, 0, 0, 1, 1, 1, 1};
    float sc[2];
    if (scales) {
        sc[0] = scales[0];
        sc[1] = scales[1];
    }
    convgemm_status_t st = convgemm_conv_ex(&d, alpha, beta, wptr, xp, yp, NULL, dtype, scales ? sc : NULL, NULL, 0,
                                            stream);
    if (st != CONVGEMM_SUCCESS) return st;
    if (!c_packed && (e = cg_launch_copy2d(yp, Nn, C, ldC, P, Nn, 4, s))) return cg_hip_fail(e, "gemm: unpacking C");
    return cg_set_error(CONVGEMM_SUCCESS, "");
}

/* ------------------------------------------------------- layout converters */

static convgemm_status_t perm(const void *src, void *dst, int d0, int d1, int d2, int p, int esize,
                              convgemm_stream_t stream) {
    if (d0 < 0 || d1 < 0 || d2 < 0 || (esize != 4 && esize != 2 && esize != 1))
        return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "layout conversion: invalid arguments");
    if ((long long)d0 * d1 * d2 == 0) return cg_set_error(CONVGEMM_SUCCESS, "");
    if (!src || !dst) return cg_set_error(CONVGEMM_ERR_INVALID_VALUE, "layout conversion: NULL pointer");
    int e = cg_launch_permute3(src, dst, d0, d1, d2, p, esize, stream);
    if (e) return cg_hip_fail(e, "layout conversion");
    return cg_set_error(CONVGEMM_SUCCESS, "");
}

convgemm_status_t convgemm_nchw_to_nhwc(const void *src, void *dst, int b, int c, int h, int w, int esize,
                                        convgemm_stream_t stream) {
    return perm(src, dst, b, c, h * w, 1, esize, stream); /* [b][c][hw] -> [b][hw][c] */
}

convgemm_status_t convgemm_nhwc_to_nchw(const void *src, void *dst, int b, int h, int w, int c, int esize,
                                        convgemm_stream_t stream) {
    return perm(src, dst, b, h * w, c, 1, esize, stream); /* [b][hw][c] -> [b][c][hw] */
}

convgemm_status_t convgemm_weights_nhwc_to_nchw(const void *src, void *dst, int c, int kh, int kw, int kn,
                                                int esize, convgemm_stream_t stream) {
    return perm(src, dst, c * kh * kw, 1, kn, 0, esize, stream); /* [K][kn] -> [kn][K] */
}

convgemm_status_t convgemm_weights_nchw_to_nhwc(const void *src, void *dst, int kn, int c, int kh, int kw,
                                                int esize, convgemm_stream_t stream) {
    return perm(src, dst, kn, 1, c * kh * kw, 0, esize, stream); /* [kn][K] -> [K][kn] */
}
