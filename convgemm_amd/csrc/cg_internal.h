/*
 * cg_internal.h -- the contract between the C host layer (host/convgemm.c)
 * and the HIP launchers (the .hip files under kernels/).  Plain C: no HIP or C++ types.
 *
 * The six convolution operations are all expressed as one implicit GEMM
 * C[M][N] = sum_k A[M][k] * B[k][N] whose operands are gathered from the
 * tensors on the fly:
 *
 *   op           M (rows)                 N        K (reduction)
 *   FWD          output pixel p           kn       (tap, c)
 *   DGRAD        input pixel of a stride  c        (valid tap, kn)
 *                residue class (ih%vs, iw%hs)
 *   WGRAD        (tap, c)                 kn       output pixel p
 *
 * tap = kx*kw + ky.  A reduction "chunk" is BK consecutive channels of one
 * tap (zero-masked past the channel count), so every K tile has a fixed tap.
 */
#ifndef CONVGEMM_AMD_CG_INTERNAL_H
#define CONVGEMM_AMD_CG_INTERNAL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum cg_opkind {
    CG_FWD_NHWC = 0,
    CG_FWD_NCHW = 1,
    CG_WGRAD_NHWC = 2,
    CG_WGRAD_NCHW = 3,
    CG_DGRAD_NHWC = 4,
    CG_DGRAD_NCHW = 5,
    CG_NUM_OPKINDS = 6
};

enum cg_elem { CG_ELEM_F32 = 0, CG_ELEM_BF16 = 1, CG_ELEM_FP8 = 2 };

typedef struct {
    int op; /* enum cg_opkind */
    int B, H, W, C, KN, KH, KW;
    int vp, hp, vs, hs, vd, hd;
    int Ho, Wo, taps;
    /* GEMM tile space (filled by the host from the chosen configuration) */
    int M, N;        /* logical rows / cols (DGRAD: rows of the largest class) */
    int mtiles, ntiles;
    int kchunks;     /* FWD/DGRAD: ceil(reduction channels / BK); WGRAD: ceil(C / BM) */
    int KT;          /* K tiles of one output tile (DGRAD: upper bound, all taps) */
    int splits;      /* split-K factor (1: no partial buffer) */
    int ncls_h, ncls_w; /* DGRAD residue classes (vs, hs); 1, 1 otherwise */
    unsigned long long cls_order; /* DGRAD with <= 16 classes: the i-th class in launch order is
                                     (cls_order >> 4i) & 15, most taps first; 0 otherwise (identity) */
    int vec;         /* reduction channels are a multiple of the 16-byte chunk (see cg_vec_ok) */
    long long P;     /* B*Ho*Wo */
    const void *A;   /* FWD/WGRAD: X;  DGRAD: dY */
    const void *Bm;  /* FWD: weights in [kn][kh][kw][c] order (repacked, or NCHW 1x1 as given);
                        DGRAD: weights in [c][kh][kw][kn] order; WGRAD: dY */
    long long a_bytes, b_bytes; /* extents of A and Bm (each < 2 GiB) */
    float *out;
    float *ws;       /* split-K partials: [splits][classes][M][N] */
    const float *bias, *mean, *inv_std, *gamma, *bnb;
    int relu;
    float alpha, beta;
    float scale;     /* product of operand scales (FP8), 1 otherwise */
    int sk;          /* stream-K workgroups (FWD/WGRAD, splits == 1); 0: off */
    int raster;      /* order of the tiles (tile_mn): 0 N panels of an M panel adjacent, 1 M panels of an
                        N panel adjacent, 2 bands of 8 M panels (M fastest within a band) */
    int sk_dp;       /* with sk: tiles 0 .. sk_dp-1 run one per workgroup (blocks 0 .. sk_dp-1), the
                        sk workgroups after them share the K steps of tiles sk_dp .. mtiles*ntiles-1 */
    int hrows, hcblk; /* halo tiles: whole output rows per M tile (Wo <= BM), or M tiles per output
                         row (Wo > BM); the other is 1 */
    int htpi, hpitch; /* halo tiles: M tiles per image, and slots per halo row (that of a full tile) */
    /* Division by the problem's divisors as a multiply and shift (cg_fill_tiles):
     * n / d == (n * m) >> s for 0 <= n < 2^31; d = Ho*Wo, Wo, C, hpitch. */
    unsigned fdm_howo, fdm_wo, fdm_c, fdm_hp;
    int fds_howo, fds_wo, fds_c, fds_hp;
    int ffd;         /* 1: the forward form of a unit-stride NHWC backward-data problem (op is CG_FWD_NHWC;
                        only the planner's problem key tells the two apart) */
    int braw;        /* FWD NHWC on a pipelined (LP) configuration: Bm holds the caller's
                        [c][kh][kw][kn] weights, read in place (no repack); 0 otherwise */
} cg_kparams;

/* Tile configurations of the implicit-GEMM kernels: a tile shape and kernel
 * (tile_tables.hpp) and the order its tiles are walked in.  Configuration
 * base + r * (number of tile shapes) is tile shape `base` in raster order r,
 * so order 0 keeps the tile shapes' own indices. */
typedef struct {
    int BM, BN, BK;
    int WM, WN;      /* waves along M and N (WM*WN = 4, or 8 for some pipelined tiles) */
    int lp;          /* > 0: a pipelined forward kernel (NHWC; NCHW for FP32) with lp LDS stages */
    int raster;      /* tile order (cg_kparams.raster) */
    int halo;        /* > 0: the halo-tile forward NHWC kernel with this many 128-byte halo slots; its M
                        tiles are whole output rows of one image, or pieces of one row (hrows, hcblk) */
} cg_tilecfg;
#define CG_NUM_RASTERS 3

int cg_num_tilecfgs(int elem);      /* tile shapes x raster orders */
int cg_num_tile_shapes(int elem);
int cg_get_tilecfg(int elem, int cfg, cg_tilecfg *out);
/* Is (elem, cfg, op, vec) instantiated? */
int cg_tilecfg_supports(int elem, int cfg, int op, int vec);

/* Launch the implicit GEMM (grid from kp); `stream` is a hipStream_t.
 * Returns 0 or a hipError_t. */
int cg_launch_igemm(int elem, int cfg, const cg_kparams *kp, void *stream);
/* Sum the split-K partials and apply the epilogue / layout store. */
int cg_launch_splitk_reduce(const cg_kparams *kp, void *stream);

/* A group of independent problems run by one launch.  Work items are
 * numbered 0..total-1; problem j owns items [start[j], start[j+1]).  Each of
 * the `grid` workgroups takes items blockIdx.x, blockIdx.x + grid, ...
 * (grid == total: one item each; grid < total: persistent workgroups).
 * plist and start are device arrays. */
typedef struct {
    const cg_kparams *plist;
    const int *start;
    int nprob;
    int total;
    int grid;
} cg_group;

/* Grouped implicit GEMM: every problem shares elem, cfg, op and vec (taken
 * from kp0, a host copy of any member); items are the problems' workgroups
 * (tiles x splits x classes). */
int cg_launch_igemm_grouped(int elem, int cfg, const cg_kparams *kp0, const cg_group *g, void *stream);
/* Grouped split-K reduction of problems sharing the operation `op`; items
 * are 256-element blocks, cg_reduce_items() per problem. */
int cg_launch_splitk_reduce_grouped(int op, const cg_group *g, void *stream);
/* The same over problems of any operations (each item switches on its problem's op). */
int cg_launch_splitk_reduce_mixed(const cg_group *g, void *stream);
long long cg_reduce_items(const cg_kparams *kp);
/* Resident workgroups per CU of a configuration (occupancy), or 0. */
int cg_igemm_occupancy(int elem, int cfg, const cg_kparams *kp0);
/* The single-launch batch kernel (kernels/single.hip): forward NHWC problems
 * with 16-byte channel chunks, each on one of a few pipelined tile shapes,
 * all in one persistent launch.  cg_single_variant: the kernel's shape index
 * of configuration cfg (any raster order), or -1 if the kernel does not run
 * that shape.  cg_launch_single: g's problems, problem j on shape var[j]
 * (device array), g->grid workgroups (a multiple of 8) walking items longest
 * problem first.  cg_single_occupancy: resident workgroups per CU. */
int cg_single_variant(int elem, int cfg);
int cg_launch_single(int elem, const cg_group *g, const int *var, void *stream);
int cg_single_occupancy(int elem);
/* Stream-K: resident workgroups per CU of the stream-K kernel of a
 * configuration (0 if it has none), and the fix-up of the tiles that several
 * workgroups share (kp->sk workgroups, partial tiles of BM x BN floats in
 * kp->ws: [workgroup][first / last piece][BM][BN]). */
int cg_igemm_sk_occupancy(int elem, int cfg, const cg_kparams *kp0);
int cg_launch_sk_fixup(const cg_kparams *kp, int BM, int BN, void *stream);

/* ---- explicit transforms (FP32) and copies ---- */
typedef struct {
    int b, h, w, c, ho, wo, kh, kw, vp, hp, vs, hs, vd, hd;
} cg_tshape;
/* nchw = 0: rows[p*ld + r] = X[b][ix][iy][c] (im2row);  nchw = 1:
 * cols[r*ld + p] = X[b][c][ix][iy] (im2col); only in-bounds taps written.
 * esize: element size, 4 (FP32) or 2 (BF16).
 * p = (b*ho + x)*wo + y, r = (c*kh + kx)*kw + ky. */
int cg_launch_im2row(const cg_tshape *s, const void *x, void *out, long long ld, int nchw, int esize, void *stream);
/* X += row2im(rows) (nchw = 0) or col2im(cols) (nchw = 1). */
int cg_launch_row2im(const cg_tshape *s, const float *in, long long ld, float *x, int nchw, void *stream);
int cg_launch_transpose_nchw(int rows, int cols, const float *in, long long ld, float beta, float *out, int kn,
                             int ho, int wo, int start_row, int start_col, void *stream);
/* dst[r*ldd + c] = src[r*lds + c] for r < rows, c < cols; esize 4, 2 or 1. */
int cg_launch_copy2d(const void *src, long long lds, void *dst, long long ldd, long long rows, long long cols,
                     int esize, void *stream);
/* dst[c*ldd + r] = src[r*lds + c] for r < rows (< 2^21), c < cols. */
int cg_launch_transpose2d(const void *src, long long lds, void *dst, long long ldd, int rows, int cols, int esize,
                          void *stream);
/* out[i] = beta*out[i] (0 if beta == 0), n elements. */
int cg_launch_scale(float *out, long long n, float beta, void *stream);
/* dst = src with its three dimensions [d0][d1][d2] permuted:
 *   perm 0: [d0][d1][d2] -> [d2][d1][d0]
 *   perm 1: [d0][d1][d2] -> [d0][d2][d1]
 *   perm 2: [d0][d1][d2] -> [d0][d1 - 1 - i][d2] (d1 reversed)
 * esize: element size in bytes (4, 2 or 1). */
/* FP32 -> BF16 (elem CG_ELEM_BF16) or FP8 e4m3fn with a per-tensor amax
 * scale (CG_ELEM_FP8; `scratch`: 8 device bytes, the scale lands at +4). */
int cg_launch_quantize(const float *src, void *dst, long long n, int elem, void *scratch, void *stream);
int cg_launch_permute3(const void *src, void *dst, int d0, int d1, int d2, int perm, int esize, void *stream);

/* Element-wise jobs of a batch run in one launch: weight repacks (kind 0, 1,
 * 3: cg_launch_permute3's perm 0, 1, 2 of [d0][d1][d2] on esize-byte
 * elements, n of them) and beta scaling of n FP32 outputs (kind 2: dst[i] =
 * beta * dst[i], 0 if beta == 0).  Job j owns items [start[j], start[j + 1]), cg_ew_items() of them
 * (32x32 transpose tiles, or blocks of CG_EW_BLOCK elements).  jobs and start
 * are device arrays. */
typedef struct {
    const void *src;
    void *dst;
    int kind, d0, d1, d2;
    long long n;
    float beta;
} cg_ewjob;
#define CG_EW_BLOCK 1024
long long cg_ew_items(const cg_ewjob *job);
int cg_launch_ew_grouped(const cg_ewjob *jobs, const int *start, int njobs, int total, int grid, int esize,
                         void *stream);

/* ---- static device scratch of the allocation-free entry point ---- */
#define CG_SCRATCH_BYTES (32 << 20)
/* Device address and size of the current device's scratch array (looked up
 * once per device with hipGetSymbolAddress).  Returns 0 or a hipError_t. */
int cg_scratch_get(void **ptr, size_t *bytes);

#ifdef __cplusplus
}
#endif

#endif
