/*
 * conv_oracle.c -- FP64 direct-convolution oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this file; it is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker.
 *
 * It restates, in double precision and by direct summation (no blocking, no
 * packing), what each public entry point of convGemm computes:
 *
 *   fwd   (sconvGemmNHWC/NCHW, trans == 'N')   src/convGemm.c:65-70, :116-119
 *   wgrad (sconvGemmNHWC/NCHW, trans != 'N')   src/convGemm.c:71-76, :120-124
 *   dgrad (sconvGemmNHWC_back/NCHW_back)       src/convGemm.c:82-97, :130-146
 *
 * Index maps (src/im2row_nhwc.c:53-58, src/im2col_nchw.c:54-58):
 *   ho = (h + 2*vpad - vdil*(kh-1) - 1)/vstr + 1            (convGemm.c:60-61)
 *   ix = vstr*x + vdil*kx - vpad,  iy = hstr*y + hdil*ky - hpad, 0 outside
 *   r  = (c*kh + kx)*kw + ky   (GEMM-K index, both layouts)
 *   p  = (b*ho + x)*wo + y     (pixel index)
 * Layouts:
 *   NHWC: X[b][h][w][c], W[c][kh][kw][kn] (offset r*kn+k), Y[b][ho][wo][kn]
 *   NCHW: X[b][c][h][w], W[kn][c][kh][kw] (offset k*K+r),  Y[b][kn][ho][wo]
 * Epilogue (im2row_nhwc.c:276-282, im2col_nchw.c:422-431), applied to
 *   t = alpha*acc + beta*y_old:  t += bias[k]; t = (t-mean[k])*inv_std[k];
 *   t = t*gamma[k] + bnbeta[k]; t = max(t, 0).  A NULL pointer means the term
 *   is absent (BN is selected by mean != NULL, as in im2col_nchw.c:484-488).
 * The reference adds y_old with coefficient 1 whenever beta != 0
 * (im2row_nhwc.c:277); alpha*acc + beta*y_old equals that for beta in {0,1}.
 *
 * Shape array s[13] = {b, h, w, c, kn, kh, kw, vpad, hpad, vstr, hstr, vdil, hdil}.
 * layout: 0 = NHWC, 1 = NCHW.
 */
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

enum { S_B, S_H, S_W, S_C, S_KN, S_KH, S_KW, S_VP, S_HP, S_VS, S_HS, S_VD, S_HD };

int oracle_out_size(int in, int pad, int dil, int k, int str) {
    return (in + 2 * pad - dil * (k - 1) - 1) / str + 1;
}

static size_t x_off(int layout, const int *s, int b, int ih, int iw, int c) {
    if (layout == 0) return (((size_t)b * s[S_H] + ih) * s[S_W] + iw) * s[S_C] + c;
    return (((size_t)b * s[S_C] + c) * s[S_H] + ih) * s[S_W] + iw;
}

static size_t w_off(int layout, const int *s, int r, int k) {
    size_t K = (size_t)s[S_C] * s[S_KH] * s[S_KW];
    if (layout == 0) return (size_t)r * s[S_KN] + k;
    return (size_t)k * K + r;
}

static size_t y_off(int layout, const int *s, int ho, int wo, int b, int x, int y, int k) {
    if (layout == 0) return (((size_t)b * ho + x) * wo + y) * s[S_KN] + k;
    return (((size_t)b * s[S_KN] + k) * ho + x) * wo + y;
}

/* Y = epi(alpha * conv(X, W) + beta * Y) */
void oracle_fwd(int layout, const int *s, double alpha, double beta,
                const double *W, const double *X, double *Y,
                const double *bias, const double *mean, const double *inv_std,
                const double *gamma, const double *bnbeta, int relu) {
    int ho = oracle_out_size(s[S_H], s[S_VP], s[S_VD], s[S_KH], s[S_VS]);
    int wo = oracle_out_size(s[S_W], s[S_HP], s[S_HD], s[S_KW], s[S_HS]);
    for (int b = 0; b < s[S_B]; b++)
        for (int x = 0; x < ho; x++)
            for (int y = 0; y < wo; y++)
                for (int k = 0; k < s[S_KN]; k++) {
                    double acc = 0.0;
                    for (int c = 0; c < s[S_C]; c++)
                        for (int kx = 0; kx < s[S_KH]; kx++) {
                            int ix = s[S_VS] * x + s[S_VD] * kx - s[S_VP];
                            if (ix < 0 || ix >= s[S_H]) continue;
                            for (int ky = 0; ky < s[S_KW]; ky++) {
                                int iy = s[S_HS] * y + s[S_HD] * ky - s[S_HP];
                                if (iy < 0 || iy >= s[S_W]) continue;
                                int r = (c * s[S_KH] + kx) * s[S_KW] + ky;
                                acc += W[w_off(layout, s, r, k)] * X[x_off(layout, s, b, ix, iy, c)];
                            }
                        }
                    size_t o = y_off(layout, s, ho, wo, b, x, y, k);
                    double t = alpha * acc;
                    if (beta != 0.0) t += beta * Y[o];
                    if (bias) t += bias[k];
                    if (mean) {
                        t = (t - mean[k]) * inv_std[k];
                        t = t * gamma[k] + bnbeta[k];
                    }
                    if (relu && t < 0.0) t = 0.0;
                    Y[o] = t;
                }
}

/* dW = alpha * im2row(X)^T dY + beta * dW   (dW in W's layout) */
void oracle_wgrad(int layout, const int *s, double alpha, double beta,
                  const double *dY, const double *X, double *dW) {
    int ho = oracle_out_size(s[S_H], s[S_VP], s[S_VD], s[S_KH], s[S_VS]);
    int wo = oracle_out_size(s[S_W], s[S_HP], s[S_HD], s[S_KW], s[S_HS]);
    for (int c = 0; c < s[S_C]; c++)
        for (int kx = 0; kx < s[S_KH]; kx++)
            for (int ky = 0; ky < s[S_KW]; ky++) {
                int r = (c * s[S_KH] + kx) * s[S_KW] + ky;
                for (int k = 0; k < s[S_KN]; k++) {
                    double acc = 0.0;
                    for (int b = 0; b < s[S_B]; b++)
                        for (int x = 0; x < ho; x++) {
                            int ix = s[S_VS] * x + s[S_VD] * kx - s[S_VP];
                            if (ix < 0 || ix >= s[S_H]) continue;
                            for (int y = 0; y < wo; y++) {
                                int iy = s[S_HS] * y + s[S_HD] * ky - s[S_HP];
                                if (iy < 0 || iy >= s[S_W]) continue;
                                acc += X[x_off(layout, s, b, ix, iy, c)] *
                                       dY[y_off(layout, s, ho, wo, b, x, y, k)];
                            }
                        }
                    size_t o = w_off(layout, s, r, k);
                    double t = alpha * acc;
                    if (beta != 0.0) t += beta * dW[o];
                    dW[o] = t;
                }
            }
}

/* dX = alpha * row2im(W^T dY) + beta * dX   (the public API uses alpha=beta=1) */
void oracle_dgrad(int layout, const int *s, double alpha, double beta,
                  const double *W, const double *dY, double *dX) {
    int ho = oracle_out_size(s[S_H], s[S_VP], s[S_VD], s[S_KH], s[S_VS]);
    int wo = oracle_out_size(s[S_W], s[S_HP], s[S_HD], s[S_KW], s[S_HS]);
    size_t nx = (size_t)s[S_B] * s[S_H] * s[S_W] * s[S_C];
    double *acc = calloc(nx ? nx : 1, sizeof(double));
    if (!acc) return;
    for (int b = 0; b < s[S_B]; b++)
        for (int x = 0; x < ho; x++)
            for (int y = 0; y < wo; y++)
                for (int c = 0; c < s[S_C]; c++)
                    for (int kx = 0; kx < s[S_KH]; kx++) {
                        int ix = s[S_VS] * x + s[S_VD] * kx - s[S_VP];
                        if (ix < 0 || ix >= s[S_H]) continue;
                        for (int ky = 0; ky < s[S_KW]; ky++) {
                            int iy = s[S_HS] * y + s[S_HD] * ky - s[S_HP];
                            if (iy < 0 || iy >= s[S_W]) continue;
                            int r = (c * s[S_KH] + kx) * s[S_KW] + ky;
                            double v = 0.0;
                            for (int k = 0; k < s[S_KN]; k++)
                                v += W[w_off(layout, s, r, k)] * dY[y_off(layout, s, ho, wo, b, x, y, k)];
                            acc[x_off(layout, s, b, ix, iy, c)] += v;
                        }
                    }
    for (size_t i = 0; i < nx; i++) {
        double t = alpha * acc[i];
        if (beta != 0.0) t += beta * dX[i];
        dX[i] = t;
    }
    free(acc);
}

/*
 * Explicit transforms, the reference's test-side oracles
 * (im2row_nhwc.c:147-229, im2col_nchw.c:146-233, :337-359).
 * rows: P x K row-major (ld K);  cols: K x P row-major (ld P).
 * Out-of-range taps are written as 0 here (the reference's tests memset the
 * buffer to 0 first, test_base.c:47-48, so the result is the same).
 */
void oracle_im2row_nhwc(const int *s, const double *X, double *rows) {
    int ho = oracle_out_size(s[S_H], s[S_VP], s[S_VD], s[S_KH], s[S_VS]);
    int wo = oracle_out_size(s[S_W], s[S_HP], s[S_HD], s[S_KW], s[S_HS]);
    size_t K = (size_t)s[S_C] * s[S_KH] * s[S_KW];
    for (int b = 0; b < s[S_B]; b++)
        for (int x = 0; x < ho; x++)
            for (int y = 0; y < wo; y++) {
                size_t p = ((size_t)b * ho + x) * wo + y;
                for (int c = 0; c < s[S_C]; c++)
                    for (int kx = 0; kx < s[S_KH]; kx++)
                        for (int ky = 0; ky < s[S_KW]; ky++) {
                            int ix = s[S_VS] * x + s[S_VD] * kx - s[S_VP];
                            int iy = s[S_HS] * y + s[S_HD] * ky - s[S_HP];
                            int r = (c * s[S_KH] + kx) * s[S_KW] + ky;
                            double v = 0.0;
                            if (ix >= 0 && ix < s[S_H] && iy >= 0 && iy < s[S_W])
                                v = X[x_off(0, s, b, ix, iy, c)];
                            rows[p * K + r] = v;
                        }
            }
}

void oracle_row2im_nhwc(const int *s, const double *rows, double *X) {
    int ho = oracle_out_size(s[S_H], s[S_VP], s[S_VD], s[S_KH], s[S_VS]);
    int wo = oracle_out_size(s[S_W], s[S_HP], s[S_HD], s[S_KW], s[S_HS]);
    size_t K = (size_t)s[S_C] * s[S_KH] * s[S_KW];
    for (int b = 0; b < s[S_B]; b++)
        for (int x = 0; x < ho; x++)
            for (int y = 0; y < wo; y++) {
                size_t p = ((size_t)b * ho + x) * wo + y;
                for (int c = 0; c < s[S_C]; c++)
                    for (int kx = 0; kx < s[S_KH]; kx++)
                        for (int ky = 0; ky < s[S_KW]; ky++) {
                            int ix = s[S_VS] * x + s[S_VD] * kx - s[S_VP];
                            int iy = s[S_HS] * y + s[S_HD] * ky - s[S_HP];
                            if (ix < 0 || ix >= s[S_H] || iy < 0 || iy >= s[S_W]) continue;
                            int r = (c * s[S_KH] + kx) * s[S_KW] + ky;
                            X[x_off(0, s, b, ix, iy, c)] += rows[p * K + r];
                        }
            }
}

void oracle_im2col_nchw(const int *s, const double *X, double *cols) {
    int ho = oracle_out_size(s[S_H], s[S_VP], s[S_VD], s[S_KH], s[S_VS]);
    int wo = oracle_out_size(s[S_W], s[S_HP], s[S_HD], s[S_KW], s[S_HS]);
    size_t P = (size_t)s[S_B] * ho * wo;
    for (int c = 0; c < s[S_C]; c++)
        for (int kx = 0; kx < s[S_KH]; kx++)
            for (int ky = 0; ky < s[S_KW]; ky++) {
                size_t r = ((size_t)c * s[S_KH] + kx) * s[S_KW] + ky;
                for (int b = 0; b < s[S_B]; b++)
                    for (int x = 0; x < ho; x++)
                        for (int y = 0; y < wo; y++) {
                            int ix = s[S_VS] * x + s[S_VD] * kx - s[S_VP];
                            int iy = s[S_HS] * y + s[S_HD] * ky - s[S_HP];
                            size_t p = ((size_t)b * ho + x) * wo + y;
                            double v = 0.0;
                            if (ix >= 0 && ix < s[S_H] && iy >= 0 && iy < s[S_W])
                                v = X[x_off(1, s, b, ix, iy, c)];
                            cols[r * P + p] = v;
                        }
            }
}

void oracle_col2im_nchw(const int *s, const double *cols, double *X) {
    int ho = oracle_out_size(s[S_H], s[S_VP], s[S_VD], s[S_KH], s[S_VS]);
    int wo = oracle_out_size(s[S_W], s[S_HP], s[S_HD], s[S_KW], s[S_HS]);
    size_t P = (size_t)s[S_B] * ho * wo;
    for (int c = 0; c < s[S_C]; c++)
        for (int kx = 0; kx < s[S_KH]; kx++)
            for (int ky = 0; ky < s[S_KW]; ky++) {
                size_t r = ((size_t)c * s[S_KH] + kx) * s[S_KW] + ky;
                for (int b = 0; b < s[S_B]; b++)
                    for (int x = 0; x < ho; x++)
                        for (int y = 0; y < wo; y++) {
                            int ix = s[S_VS] * x + s[S_VD] * kx - s[S_VP];
                            int iy = s[S_HS] * y + s[S_HD] * ky - s[S_HP];
                            if (ix < 0 || ix >= s[S_H] || iy < 0 || iy >= s[S_W]) continue;
                            size_t p = ((size_t)b * ho + x) * wo + y;
                            X[x_off(1, s, b, ix, iy, c)] += cols[r * P + p];
                        }
            }
}

/*
 * Swap the first two dimensions of a [d0][d1][inner] tensor:
 * out[j][i][:] = in[i][j][:] (+ beta*out).  transpose_nchw
 * (im2col_nchw.c:206-233) is this with (d0, d1) = (kn, b), inner = ho*wo.
 */
void oracle_swap01(int d0, int d1, int inner, double beta, const double *in, double *out) {
    for (int i = 0; i < d0; i++)
        for (int j = 0; j < d1; j++)
            for (int t = 0; t < inner; t++) {
                size_t src = ((size_t)i * d1 + j) * inner + t;
                size_t dst = ((size_t)j * d0 + i) * inner + t;
                out[dst] = (beta == 0.0) ? in[src] : beta * out[dst] + in[src];
            }
}
