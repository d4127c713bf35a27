/*
 * convgemm_port.c -- FP32 CPU restatement of convGemm's implicit-GEMM algorithm.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  It is the timed CPU baseline of
 * bench.py ("kind": "port") and a second, FP32 checker for the tests.  Nothing
 * in the product library links or calls it.
 *
 * Algorithm restated (own code, same steps):
 *   - BLIS five-loop GEMM C = beta*C + alpha*op(A)*op(B):
 *       jc (NC) -> pc (KC, betaI = beta on the first block only, `last` flag)
 *       -> pack B block into NR-column panels -> ic (MC) -> pack A block into
 *       MR-row panels -> jr (NR) -> ir (MR) -> micro-kernel -> post-process.
 *     src/gemm_blis_B3A2C0_orig.c:55-110 (parallel jr) and
 *     src/gemm_blis_B3A2C0.c:55-115 (parallel ic, per-thread A panel).
 *   - quick return / alpha==0 (B3A2C0_orig.c:45-47, quick_gemm.h:1-30).
 *   - implicit im2row / im2col packing: the convolution patches are gathered
 *     straight into the packed panels (im2row_nhwc.c:28-145,
 *     im2col_nchw.c:29-144, :238-335).
 *   - post-processing on each MR x NR tile: plain beta write (sxpbyM,
 *     gemm_blis.c:147-161), the fused bias/BN/ReLU epilogue
 *     (im2row_nhwc.c:269-372), the transposing NCHW epilogue
 *     (im2col_nchw.c:397-493) and the atomic row2im/col2im scatter of the
 *     backward-data pass (im2row_nhwc.c:231-267, im2col_nchw.c:361-395).
 *   - per-entry-point orientation of the GEMM (src/convGemm.c:48-146).
 * The micro-kernel is portable C (the reference uses BLIS's assembly
 * micro-kernel, which is not available here), so absolute CPU speed is lower
 * than the reference's; every report labels this a restatement.
 */
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define PMR 12
#define PNR 16
#define PMC 120
#define PKC 256
#define PNC 3072

enum { OP_FWD = 0, OP_WGRAD = 1, OP_DGRAD = 2 };
enum { L_NHWC = 0, L_NCHW = 1 };
enum { POST_NONE, POST_FWD_NHWC, POST_FWD_NCHW, POST_ROW2IM, POST_COL2IM };

typedef struct {
    int op, layout;
    int b, h, w, c, kn, kh, kw, vp, hp, vs, hs, vd, hd, ho, wo;
    const float *in, *x;   /* W (fwd, dgrad) or dY (wgrad); X (fwd, wgrad) or dY (dgrad) */
    float *out;
    const float *bias, *mean, *inv_std, *gamma, *bnb;
    bool relu;
    int post;
} pctx;

static int g_threads = 0;

void port_set_threads(int n) { g_threads = n; }

int port_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

/* ---------------- index decomposition helpers ---------------- */

/* pixel p = (b*ho + x)*wo + y  ->  b, x*vs - vp, y*hs - hp */
static inline void split_pixel(const pctx *t, long p, int *b, int *px, int *py, int *pix) {
    int y = (int)(p % t->wo);
    long q = p / t->wo;
    int x = (int)(q % t->ho);
    *b = (int)(q / t->ho);
    *px = x * t->vs - t->vp;
    *py = y * t->hs - t->hp;
    *pix = x * t->wo + y;
}

/* GEMM-K index r = (c*kh + kx)*kw + ky  ->  c, kx*vd, ky*hd */
static inline void split_tap(const pctx *t, int r, int *c, int *rx, int *ry) {
    int ky = r % t->kw;
    int q = r / t->kw;
    *rx = (q % t->kh) * t->vd;
    *ry = ky * t->hd;
    *c = q / t->kh;
}

/* X gather for patch element (p, r) of im2row (NHWC) / im2col (NCHW) */
static inline float x_gather(const pctx *t, int b, int px, int py, int c, int rx, int ry) {
    int ih = px + rx, iw = py + ry;
    if (ih < 0 || ih >= t->h || iw < 0 || iw >= t->w) return 0.0f;
    if (t->layout == L_NHWC)
        return t->x[(((long)b * t->h + ih) * t->w + iw) * t->c + c];
    return t->x[(((long)b * t->c + c) * t->h + ih) * t->w + iw];
}

/* ---------------- operand element access ---------------- */
/* The GEMM for each entry point is the one src/convGemm.c sets up:
 *  fwd NHWC  m=kn n=P k=K  A(k,r)=W[r*kn+k]      B(r,p)=im2row
 *  fwd NCHW  m=P  n=kn k=K A(p,r)=im2col         B(r,k)=W[k*K+r]
 *  wgrad NHWC m=kn n=K k=P A(k,p)=dY[p*kn+k]     B(p,r)=im2row
 *  wgrad NCHW m=K n=kn k=P A(r,p)=im2col         B(p,k)=dY_nchw
 *  dgrad NHWC m=K n=P k=kn A(r,j)=W[r*kn+j]      B(j,p)=dY[p*kn+j]
 *  dgrad NCHW m=P n=K k=kn A(p,j)=dY_nchw        B(j,r)=W[j*K+r]          */

static inline float dy_nchw(const pctx *t, const float *dy, long p, int k) {
    int b, px, py, pix;
    split_pixel(t, p, &b, &px, &py, &pix);
    return dy[((long)b * t->kn + k) * t->ho * t->wo + pix];
}

/* Pack an mc x kc block of A starting at (i0, k0) into MR-row panels:
 * Ac[panel][kk][ii] (zero fill past mc, as pack_RB, gemm_blis.c:40-88). */
static void pack_a(const pctx *t, int i0, int k0, int mc, int kc, float *Ac, bool parallel) {
    int npanels = (mc + PMR - 1) / PMR;
    /* parallel: panels are split over the threads of the enclosing team
     * (the caller places the barriers); otherwise one thread packs all. */
    for (int pn = 0; pn < npanels; pn++) {
        if (parallel && (pn % omp_get_num_threads()) != omp_get_thread_num()) continue;
        float *dst = Ac + (long)pn * PMR * kc;
        for (int ii = 0; ii < PMR; ii++) {
            int i = i0 + pn * PMR + ii;
            bool valid_row = (pn * PMR + ii) < mc;
            if (!valid_row) {
                for (int kk = 0; kk < kc; kk++) dst[kk * PMR + ii] = 0.0f;
                continue;
            }
            switch (t->op * 2 + t->layout) {
            case OP_FWD * 2 + L_NHWC:   /* A(k, r) = W[r*kn + k] */
            case OP_DGRAD * 2 + L_NHWC: /* A(r, j) = W[r*kn + j] */
                if (t->op == OP_FWD)
                    for (int kk = 0; kk < kc; kk++) dst[kk * PMR + ii] = t->in[(long)(k0 + kk) * t->kn + i];
                else
                    for (int kk = 0; kk < kc; kk++) dst[kk * PMR + ii] = t->in[(long)i * t->kn + k0 + kk];
                break;
            case OP_FWD * 2 + L_NCHW: { /* A(p, r) = im2col */
                int b, px, py, pix;
                split_pixel(t, i, &b, &px, &py, &pix);
                for (int kk = 0; kk < kc; kk++) {
                    int c, rx, ry;
                    split_tap(t, k0 + kk, &c, &rx, &ry);
                    dst[kk * PMR + ii] = x_gather(t, b, px, py, c, rx, ry);
                }
                break;
            }
            case OP_WGRAD * 2 + L_NHWC: /* A(k, p) = dY[p*kn + k] */
                for (int kk = 0; kk < kc; kk++) dst[kk * PMR + ii] = t->in[(long)(k0 + kk) * t->kn + i];
                break;
            case OP_WGRAD * 2 + L_NCHW: { /* A(r, p) = im2col */
                int c, rx, ry;
                split_tap(t, i, &c, &rx, &ry);
                for (int kk = 0; kk < kc; kk++) {
                    int b, px, py, pix;
                    split_pixel(t, k0 + kk, &b, &px, &py, &pix);
                    dst[kk * PMR + ii] = x_gather(t, b, px, py, c, rx, ry);
                }
                break;
            }
            case OP_DGRAD * 2 + L_NCHW: /* A(p, j) = dY_nchw */
                for (int kk = 0; kk < kc; kk++) dst[kk * PMR + ii] = dy_nchw(t, t->x, i, k0 + kk);
                break;
            }
        }
    }
}

/* Pack a kc x nc block of B starting at (k0, j0) into NR-column panels:
 * Bc[panel][kk][jj] (zero fill past nc, as pack_CB, gemm_blis.c:93-142). */
static void pack_b(const pctx *t, int k0, int j0, int kc, int nc, float *Bc) {
    long K = (long)t->c * t->kh * t->kw;
    int npanels = (nc + PNR - 1) / PNR;
#pragma omp for schedule(static)
    for (int pn = 0; pn < npanels; pn++) {
        float *dst = Bc + (long)pn * PNR * kc;
        for (int jj = 0; jj < PNR; jj++) {
            int j = j0 + pn * PNR + jj;
            if (pn * PNR + jj >= nc) {
                for (int kk = 0; kk < kc; kk++) dst[kk * PNR + jj] = 0.0f;
                continue;
            }
            switch (t->op * 2 + t->layout) {
            case OP_FWD * 2 + L_NHWC: { /* B(r, p) = im2row */
                int b, px, py, pix;
                split_pixel(t, j, &b, &px, &py, &pix);
                for (int kk = 0; kk < kc; kk++) {
                    int c, rx, ry;
                    split_tap(t, k0 + kk, &c, &rx, &ry);
                    dst[kk * PNR + jj] = x_gather(t, b, px, py, c, rx, ry);
                }
                break;
            }
            case OP_FWD * 2 + L_NCHW: /* B(r, k) = W[k*K + r] */
                for (int kk = 0; kk < kc; kk++) dst[kk * PNR + jj] = t->in[(long)j * K + k0 + kk];
                break;
            case OP_WGRAD * 2 + L_NHWC: { /* B(p, r) = im2row */
                int c, rx, ry;
                split_tap(t, j, &c, &rx, &ry);
                for (int kk = 0; kk < kc; kk++) {
                    int b, px, py, pix;
                    split_pixel(t, k0 + kk, &b, &px, &py, &pix);
                    dst[kk * PNR + jj] = x_gather(t, b, px, py, c, rx, ry);
                }
                break;
            }
            case OP_WGRAD * 2 + L_NCHW: /* B(p, k) = dY_nchw */
                for (int kk = 0; kk < kc; kk++) dst[kk * PNR + jj] = dy_nchw(t, t->in, k0 + kk, j);
                break;
            case OP_DGRAD * 2 + L_NHWC: /* B(j, p) = dY[p*kn + j] */
                for (int kk = 0; kk < kc; kk++) dst[kk * PNR + jj] = t->x[(long)j * t->kn + k0 + kk];
                break;
            case OP_DGRAD * 2 + L_NCHW: /* B(j, r) = W[j*K + r] */
                for (int kk = 0; kk < kc; kk++) dst[kk * PNR + jj] = t->in[(long)(k0 + kk) * K + j];
                break;
            }
        }
    }
}

/* MR x NR micro-kernel: c[i][j] = alpha * sum_p a[p][i] * b[p][j] */
__attribute__((target_clones("avx512f", "avx2", "default")))
static void ukr(int kc, float alpha, const float *restrict a, const float *restrict b, float *restrict c) {
    float acc[PMR][PNR];
    for (int i = 0; i < PMR; i++)
        for (int j = 0; j < PNR; j++) acc[i][j] = 0.0f;
    for (int p = 0; p < kc; p++) {
        const float *ap = a + p * PMR;
        const float *bp = b + p * PNR;
        for (int i = 0; i < PMR; i++) {
            float ai = ap[i];
#pragma omp simd
            for (int j = 0; j < PNR; j++) acc[i][j] += ai * bp[j];
        }
    }
    for (int i = 0; i < PMR; i++)
        for (int j = 0; j < PNR; j++) c[i * PNR + j] = alpha * acc[i][j];
}

/* Post-process one mr x nr tile whose top-left GEMM coordinate is (i0, j0).
 * ct is the alpha-scaled product of this K block, row-major PMR x PNR.      */
static void post_tile(const pctx *t, int i0, int j0, int mr, int nr, const float *ct,
                      float betaI, bool last, float *C, long ldC) {
    switch (t->post) {
    case POST_NONE: /* C column-major, sxpbyM semantics (gemm_blis.c:147-161) */
        for (int j = 0; j < nr; j++)
            for (int i = 0; i < mr; i++) {
                float *cp = &C[(long)(j0 + j) * ldC + i0 + i];
                float v = ct[i * PNR + j];
                if (betaI == 0.0f) *cp = v;
                else if (betaI == 1.0f) *cp += v;
                else *cp = betaI * *cp + v;
            }
        break;
    case POST_FWD_NHWC: /* C = Y, kn x P col-major; rows i = channel k */
        for (int j = 0; j < nr; j++)
            for (int i = 0; i < mr; i++) {
                int k = i0 + i;
                float *cp = &C[(long)(j0 + j) * ldC + k];
                float v = ct[i * PNR + j];
                if (!last) {
                    if (betaI == 0.0f) *cp = v;
                    else if (betaI == 1.0f) *cp += v;
                    else *cp = betaI * *cp + v;
                    continue;
                }
                if (betaI != 0.0f) v += betaI * *cp;
                if (t->bias) v += t->bias[k];
                if (t->mean) {
                    v = (v - t->mean[k]) * t->inv_std[k];
                    v = v * t->gamma[k] + t->bnb[k];
                }
                if (t->relu && v < 0.0f) v = 0.0f;
                *cp = v;
            }
        break;
    case POST_FWD_NCHW: /* rows i = pixel p, cols j = channel k; Y[b][k][pix] */
        for (int i = 0; i < mr; i++) {
            int b, px, py, pix;
            split_pixel(t, i0 + i, &b, &px, &py, &pix);
            for (int j = 0; j < nr; j++) {
                int k = j0 + j;
                float *cp = &t->out[((long)b * t->kn + k) * t->ho * t->wo + pix];
                float v = ct[i * PNR + j];
                if (betaI != 0.0f) v += betaI * *cp;
                if (last) {
                    if (t->bias) v += t->bias[k];
                    if (t->mean) {
                        v = (v - t->mean[k]) * t->inv_std[k];
                        v = v * t->gamma[k] + t->bnb[k];
                    }
                    if (t->relu && v < 0.0f) v = 0.0f;
                }
                *cp = v;
            }
        }
        break;
    case POST_ROW2IM: /* rows i = r, cols j = pixel; dX NHWC += (atomic) */
        for (int j = 0; j < nr; j++) {
            int b, px, py, pix;
            split_pixel(t, j0 + j, &b, &px, &py, &pix);
            for (int i = 0; i < mr; i++) {
                int c, rx, ry;
                split_tap(t, i0 + i, &c, &rx, &ry);
                int ih = px + rx, iw = py + ry;
                if (ih < 0 || ih >= t->h || iw < 0 || iw >= t->w) continue;
                float *dp = &t->out[(((long)b * t->h + ih) * t->w + iw) * t->c + c];
#pragma omp atomic
                *dp += ct[i * PNR + j];
            }
        }
        break;
    case POST_COL2IM: /* rows i = pixel, cols j = r; dX NCHW += (atomic) */
        for (int i = 0; i < mr; i++) {
            int b, px, py, pix;
            split_pixel(t, i0 + i, &b, &px, &py, &pix);
            for (int j = 0; j < nr; j++) {
                int c, rx, ry;
                split_tap(t, j0 + j, &c, &rx, &ry);
                int ih = px + rx, iw = py + ry;
                if (ih < 0 || ih >= t->h || iw < 0 || iw >= t->w) continue;
                float *dp = &t->out[(((long)b * t->c + c) * t->h + ih) * t->w + iw];
#pragma omp atomic
                *dp += ct[i * PNR + j];
            }
        }
        break;
    }
}

/* Five-loop GEMM with implicit packing.  ic_parallel selects the B3A2C0
 * order (parallel ic, one A panel per thread) instead of B3A2C0_orig
 * (parallel jr, shared A panel). */
static void port_gemm(const pctx *t, int m, int n, int k, float alpha, float beta,
                      float *C, long ldC, bool ic_parallel) {
    if (m <= 0 || n <= 0 || ((alpha == 0.0f || k == 0) && beta == 1.0f)) return;
    if (alpha == 0.0f) { /* quick_gemm.h: C = beta*C (0 if beta == 0) */
        if (t->post == POST_ROW2IM || t->post == POST_COL2IM) {
            /* dX is the output tensor, not an m x n matrix */
            long nx = (long)t->b * t->h * t->w * t->c;
            for (long i = 0; i < nx; i++) C[i] = beta == 0.0f ? 0.0f : beta * C[i];
            return;
        }
        for (long j = 0; j < n; j++)
            for (long i = 0; i < m; i++)
                C[j * ldC + i] = beta == 0.0f ? 0.0f : beta * C[j * ldC + i];
        return;
    }
    int nthr = port_get_threads();
    float *Bc = aligned_alloc(64, sizeof(float) * (size_t)PKC * (PNC + PNR));
    float *Ac = aligned_alloc(64, sizeof(float) * (size_t)nthr * PKC * (PMC + PMR));
    if (!Ac || !Bc) { free(Ac); free(Bc); return; }

#pragma omp parallel num_threads(nthr)
    {
        float ct[PMR * PNR];
        int tid = omp_get_thread_num();
        for (int jc = 0; jc < n; jc += PNC) {
            int nc = n - jc < PNC ? n - jc : PNC;
            for (int pc = 0; pc < k; pc += PKC) {
                int kc = k - pc < PKC ? k - pc : PKC;
                bool last = pc + PKC >= k;
                float betaI = pc == 0 ? beta : 1.0f;
                pack_b(t, pc, jc, kc, nc, Bc); /* implicit barrier */
                if (!ic_parallel) {
                    for (int ic = 0; ic < m; ic += PMC) {
                        int mc = m - ic < PMC ? m - ic : PMC;
#pragma omp barrier
                        pack_a(t, ic, pc, mc, kc, Ac, true);
#pragma omp barrier
#pragma omp for schedule(static)
                        for (int jr = 0; jr < nc; jr += PNR) {
                            int nr = nc - jr < PNR ? nc - jr : PNR;
                            for (int ir = 0; ir < mc; ir += PMR) {
                                int mr = mc - ir < PMR ? mc - ir : PMR;
                                ukr(kc, alpha, Ac + (long)ir * kc, Bc + (long)jr * kc, ct);
                                post_tile(t, ic + ir, jc + jr, mr, nr, ct, betaI, last, C, ldC);
                            }
                        }
                    }
                } else {
                    float *Acl = Ac + (long)tid * PKC * (PMC + PMR);
#pragma omp for schedule(dynamic, 1)
                    for (int ic = 0; ic < m; ic += PMC) {
                        int mc = m - ic < PMC ? m - ic : PMC;
                        pack_a(t, ic, pc, mc, kc, Acl, false);
                        for (int jr = 0; jr < nc; jr += PNR) {
                            int nr = nc - jr < PNR ? nc - jr : PNR;
                            for (int ir = 0; ir < mc; ir += PMR) {
                                int mr = mc - ir < PMR ? mc - ir : PMR;
                                ukr(kc, alpha, Acl + (long)ir * kc, Bc + (long)jr * kc, ct);
                                post_tile(t, ic + ir, jc + jr, mr, nr, ct, betaI, last, C, ldC);
                            }
                        }
                    }
                }
#pragma omp barrier
            }
        }
    }
    free(Ac);
    free(Bc);
}

static void fill_ctx(pctx *t, int op, int layout, int b, int h, int w, int c, int kn, int kh, int kw,
                     int vp, int hp, int vs, int hs, int vd, int hd) {
    memset(t, 0, sizeof(*t));
    t->op = op; t->layout = layout;
    t->b = b; t->h = h; t->w = w; t->c = c; t->kn = kn; t->kh = kh; t->kw = kw;
    t->vp = vp; t->hp = hp; t->vs = vs; t->hs = hs; t->vd = vd; t->hd = hd;
    t->ho = (h + 2 * vp - vd * (kh - 1) - 1) / vs + 1;
    t->wo = (w + 2 * hp - hd * (kw - 1) - 1) / hs + 1;
}

/* Same argument order and meaning as sconvGemmNHWC (src/convGemm.h:22-33). */
void port_sconvGemmNHWC(char trans, int b, int h, int w, int c, int kn, int kh, int kw,
                        int vpadding, int hpadding, int vstride, int hstride, int vdilation, int hdilation,
                        const float *in, const float *x, float *out, const float *bias_vector,
                        const float *bn_running_mean, const float *bn_inv_std,
                        const float *bn_gamma, const float *bn_beta, bool relu) {
    pctx t;
    fill_ctx(&t, trans == 'N' ? OP_FWD : OP_WGRAD, L_NHWC, b, h, w, c, kn, kh, kw,
             vpadding, hpadding, vstride, hstride, vdilation, hdilation);
    t.in = in; t.x = x; t.out = out;
    int P = b * t.ho * t.wo, K = c * kh * kw;
    if (trans == 'N') {
        t.bias = bias_vector; t.mean = bn_running_mean; t.inv_std = bn_inv_std;
        t.gamma = bn_gamma; t.bnb = bn_beta; t.relu = relu;
        t.post = (bias_vector || bn_running_mean || relu) ? POST_FWD_NHWC : POST_NONE;
        port_gemm(&t, kn, P, K, 1.0f, 0.0f, out, kn, false);
    } else {
        t.post = POST_NONE;
        port_gemm(&t, kn, K, P, 1.0f, 0.0f, out, kn, false);
    }
}

/* Same argument order and meaning as sconvGemmNHWC_back (src/convGemm.h:35-42). */
void port_sconvGemmNHWC_back(int b, int h, int w, int c, int kn, int kh, int kw,
                             int vstride, int hstride, int vpadding, int hpadding, int vdilation, int hdilation,
                             const float *weights, const float *dy, float *dx) {
    pctx t;
    fill_ctx(&t, OP_DGRAD, L_NHWC, b, h, w, c, kn, kh, kw, vpadding, hpadding, vstride, hstride,
             vdilation, hdilation);
    t.in = weights; t.x = dy; t.out = dx; t.post = POST_ROW2IM;
    port_gemm(&t, c * kh * kw, b * t.ho * t.wo, kn, 1.0f, 1.0f, dx, c * kh * kw, false);
}

/* Same argument order and meaning as sconvGemmNCHW (src/convGemm.h:44-55). */
void port_sconvGemmNCHW(char trans, int b, int c, int h, int w, int kn, int kh, int kw,
                        int vpadding, int hpadding, int vstride, int hstride, int vdilation, int hdilation,
                        const float *in, const float *x, float *out, const float *bias_vector,
                        const float *bn_running_mean, const float *bn_inv_std,
                        const float *bn_gamma, const float *bn_beta, bool relu) {
    pctx t;
    fill_ctx(&t, trans == 'N' ? OP_FWD : OP_WGRAD, L_NCHW, b, h, w, c, kn, kh, kw,
             vpadding, hpadding, vstride, hstride, vdilation, hdilation);
    t.in = in; t.x = x; t.out = out;
    int P = b * t.ho * t.wo, K = c * kh * kw;
    if (trans == 'N') {
        t.bias = bias_vector; t.mean = bn_running_mean; t.inv_std = bn_inv_std;
        t.gamma = bn_gamma; t.bnb = bn_beta; t.relu = relu;
        t.post = POST_FWD_NCHW; /* always installed: it performs the NCHW store */
        port_gemm(&t, P, kn, K, 1.0f, 0.0f, out, P, true);
    } else {
        t.post = POST_NONE;
        port_gemm(&t, K, kn, P, 1.0f, 0.0f, out, K, true);
    }
}

/* Same argument order and meaning as sconvGemmNCHW_back (src/convGemm.h:57-64). */
void port_sconvGemmNCHW_back(int b, int c, int h, int w, int kn, int kh, int kw,
                             int vstride, int hstride, int vpadding, int hpadding, int vdilation, int hdilation,
                             const float *weights, const float *dy, float *dx) {
    pctx t;
    fill_ctx(&t, OP_DGRAD, L_NCHW, b, h, w, c, kn, kh, kw, vpadding, hpadding, vstride, hstride,
             vdilation, hdilation);
    t.in = weights; t.x = dy; t.out = dx; t.post = POST_COL2IM;
    port_gemm(&t, b * t.ho * t.wo, c * kh * kw, kn, 1.0f, 1.0f, dx, b * t.ho * t.wo, false);
}
