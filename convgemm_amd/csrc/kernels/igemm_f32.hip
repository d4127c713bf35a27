// igemm_f32.hip -- FP32 implicit-GEMM convolution on gfx950 MFMA.
//
// One kernel template serves the six operations of convGemm
// (src/convGemm.c:48-146).  The reference packs im2row/im2col panels on the
// CPU (src/im2row_nhwc.c:28-145, src/im2col_nchw.c:29-144, :238-335) and
// runs a BLIS micro-kernel on them; here each workgroup gathers its
// convolution patches straight from HBM into an LDS double buffer and runs
// the contraction on v_mfma_f32_32x32x2f32.  The reference's post-process
// callbacks (bias/BN/ReLU epilogue, NCHW transposed store, row2im/col2im
// scatter) become the write-back of the accumulators: backward data is
// computed in gather form (one writer per dX element, no atomics), split by
// stride residue class so that no multiply touches a structural zero.
//
// Tile: BM x BN outputs per 256-thread workgroup, BK reduction elements per
// K step, 4 waves arranged WM x WN, each wave owning (BM/WM) x (BN/WN) as
// 32x32 MFMA blocks.  LDS rows are K-contiguous with a 4-float pad.
#include "igemm_common.hpp"

namespace cg {

constexpr int NT = 256;

enum LoadKind { LK_KV = 0, LK_RS = 1, LK_KS = 2 };

template <int OP> constexpr bool is_fwd() { return OP == CG_FWD_NHWC || OP == CG_FWD_NCHW; }
template <int OP> constexpr bool is_dgrad() { return OP == CG_DGRAD_NHWC || OP == CG_DGRAD_NCHW; }
template <int OP> constexpr bool is_wgrad() { return OP == CG_WGRAD_NHWC || OP == CG_WGRAD_NCHW; }
// NCHW outputs are contiguous along M: put M on the MFMA columns (lanes).
template <int OP> constexpr bool swap_out() {
    return OP == CG_FWD_NCHW || OP == CG_DGRAD_NCHW || OP == CG_WGRAD_NCHW;
}
// How each operand tile is read from HBM:
//   KV: contiguous along K, 16-byte vectors (4 K values per thread)
//   KS: contiguous along K, 4 scalar loads (unaligned runs)
//   RS: contiguous along the row dimension (lanes walk rows), 4 K values/thread
template <int OP> constexpr int a_kind() {
    return OP == CG_FWD_NHWC ? LK_KV : OP == CG_FWD_NCHW ? LK_RS : OP == CG_DGRAD_NHWC ? LK_KV
           : OP == CG_DGRAD_NCHW ? LK_RS : OP == CG_WGRAD_NHWC ? LK_RS : LK_KS;
}
template <int OP> constexpr int b_kind() {
    return is_fwd<OP>() ? LK_RS : is_dgrad<OP>() ? LK_KV : OP == CG_WGRAD_NHWC ? LK_RS : LK_KS;
}

template <int KIND, int R, int BK> constexpr bool loader_ok() {
    constexpr int KQ = BK / 4;
    if (R * KQ < NT) return false;
    if (KIND == LK_RS) return NT % R == 0;
    return (R * KQ) % NT == 0;
}

// Thread -> (row, k-quad) of a loader.
template <int KIND, int R, int BK>
__device__ inline int ld_row(int tid, int i) {
    constexpr int KQ = BK / 4;
    if constexpr (KIND == LK_RS) return tid % R;
    else return tid / KQ + i * (NT / KQ);
}
template <int KIND, int R, int BK>
__device__ inline int ld_kq(int tid, int i) {
    constexpr int KQ = BK / 4;
    if constexpr (KIND == LK_RS) return tid / R + i * (NT / R);
    else return tid % KQ;
}

constexpr int FAR = -(1 << 29);  // row coordinate that fails every bounds test

template <int BM, int BN, int BK, int WM, int WN, int OP>
__global__ void __launch_bounds__(NT) igemm_f32_kernel(const cg_kparams p) {
    constexpr int PITCH = BK + 4;
    constexpr int KQ = BK / 4;
    constexpr int PA = BM * KQ / NT;
    constexpr int PB = BN * KQ / NT;
    constexpr int AK = a_kind<OP>();
    constexpr int BKIND = b_kind<OP>();
    constexpr bool FW = is_fwd<OP>(), DG = is_dgrad<OP>(), WG = is_wgrad<OP>();
    constexpr bool SW = swap_out<OP>();
    constexpr int NRA = (AK == LK_RS) ? 1 : PA;  // distinct A rows per thread
    constexpr int NRB = (BKIND == LK_RS) ? 1 : PB;
    static_assert(loader_ok<AK, BM, BK>() && loader_ok<BKIND, BN, BK>(), "loader geometry");
    static_assert(BK % 16 == 0, "BK");
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int ER = SW ? WTN : WTM, EC = SW ? WTM : WTN;
    constexpr int TR = ER / 32, TC = EC / 32;
    static_assert(TR >= 1 && TC >= 1 && ER % 32 == 0 && EC % 32 == 0, "wave tile");

    __shared__ __attribute__((aligned(16))) float sA[2 * BM * PITCH];
    __shared__ __attribute__((aligned(16))) float sB[2 * BN * PITCH];
    __shared__ int s_kx[64], s_ky[64], s_nk[2];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int mt = blockIdx.x % p.mtiles, nt = blockIdx.x / p.mtiles;
    const int cls = blockIdx.y, split = blockIdx.z;
    const Ctx x = make_ctx(p, cls);
    if (!WG && mt * BM >= x.Mrows) return;  // uniform: DGRAD classes smaller than the grid

    const float *__restrict__ Asrc = (const float *)p.A;
    const float *__restrict__ Bsrc = (const float *)p.Bm;
    const int HoWo = p.Ho * p.Wo;

    // ---- reduction range of this block ----
    int nkx = p.KH, nky = p.KW;
    if constexpr (DG) {
        if (tid == 0) {
            int n = 0;
            for (int kx = 0; kx < p.KH; kx++)
                if ((x.rh + p.vp - kx * p.vd) % p.vs == 0) s_kx[n++] = kx;
            s_nk[0] = n;
            n = 0;
            for (int ky = 0; ky < p.KW; ky++)
                if ((x.rw + p.hp - ky * p.hd) % p.hs == 0) s_ky[n++] = ky;
            s_nk[1] = n;
        }
        __syncthreads();
        nkx = s_nk[0];
        nky = s_nk[1];
    }
    const int KT = WG ? p.KT : nkx * nky * p.kchunks;
    const int per = (KT + p.splits - 1) / p.splits;
    const int kt_begin = min(KT, split * per);
    const int kt_end = min(KT, kt_begin + per);

    // WGRAD: this M tile is (tap, c0 .. c0+BM)
    int wg_tap = 0, wg_c0 = 0, wg_kx = 0, wg_ky = 0;
    if constexpr (WG) {
        wg_tap = mt / p.kchunks;
        wg_c0 = (mt - wg_tap * p.kchunks) * BM;
        wg_kx = wg_tap / p.KW;
        wg_ky = wg_tap - wg_kx * p.KW;
    }

    // ---- per-thread row state ----
    int a_b[NRA], a_h[NRA], a_w[NRA];  // FWD: b, x*vs-vp, y*hs-hp; DGRAD: b, i', j'; WGRAD: -, c, -
#pragma unroll
    for (int i = 0; i < NRA; i++) {
        int row = ld_row<AK, BM, BK>(tid, i);
        if constexpr (FW) {
            int m = mt * BM + row;
            a_b[i] = 0;
            a_h[i] = FAR;
            a_w[i] = FAR;
            if (m < (int)p.P) {
                int b = fdiv(m, x.howo);
                int pix = m - b * HoWo;
                int xo = fdiv(pix, x.wo);
                a_b[i] = b;
                a_h[i] = xo * p.vs - p.vp;
                a_w[i] = (pix - xo * p.Wo) * p.hs - p.hp;
            }
        } else if constexpr (DG) {
            int q = mt * BM + row;
            a_b[i] = 0;
            a_h[i] = FAR;
            a_w[i] = FAR;
            if (q < x.Mrows) {
                int b = fdiv(q, x.cls_hw);
                int rem = q - b * (x.Hc * x.Wc);
                int ii = fdiv(rem, x.cls_w);
                a_b[i] = b;
                a_h[i] = ii;
                a_w[i] = rem - ii * x.Wc;
            }
        } else {
            int c = wg_c0 + row;
            a_b[i] = 0;
            a_h[i] = c < p.C ? c : -1;
            a_w[i] = 0;
        }
    }
    int b_row[NRB];  // FWD/WGRAD: kn index (or -1); DGRAD: c index (or -1)
#pragma unroll
    for (int i = 0; i < NRB; i++) {
        int r = nt * BN + ld_row<BKIND, BN, BK>(tid, i);
        b_row[i] = r < (DG ? p.C : p.KN) ? r : -1;
    }

    f32x4 ra[PA], rb[PB];

    // ---- gather one K tile (global -> registers) ----
    auto load_tile = [&](int kt) {
        if constexpr (!WG) {
            const int ti = kt / p.kchunks;
            const int r0 = (kt - ti * p.kchunks) * BK;
            int kx, ky;
            if constexpr (DG) {
                int vx = ti / nky;
                kx = s_kx[vx];
                ky = s_ky[ti - vx * nky];
            } else {
                kx = ti / p.KW;
                ky = ti - kx * p.KW;
            }
            const int tap = kx * p.KW + ky;
            // A operand
            if constexpr (OP == CG_FWD_NHWC || OP == CG_DGRAD_NHWC) {
                const int kq = ld_kq<AK, BM, BK>(tid, 0);
                const int c = r0 + kq * 4;
                const int cred = FW ? p.C : p.KN;
                int dh, dw, lim_h, lim_w;
                if constexpr (FW) {
                    dh = kx * p.vd; dw = ky * p.hd; lim_h = p.H; lim_w = p.W;
                } else {
                    dh = (x.rh + p.vp - kx * p.vd) / p.vs;
                    dw = (x.rw + p.hp - ky * p.hd) / p.hs;
                    lim_h = p.Ho; lim_w = p.Wo;
                }
#pragma unroll
                for (int i = 0; i < PA; i++) {
                    const int ih = a_h[i] + dh, iw = a_w[i] + dw;
                    const bool okp = (unsigned)ih < (unsigned)lim_h && (unsigned)iw < (unsigned)lim_w;
                    const long long base = (((long long)a_b[i] * lim_h + ih) * lim_w + iw) * cred + c;
                    if (p.vecA) {
                        ra[i] = (okp && c < cred) ? *(const f32x4 *)(Asrc + base) : f32x4{0.f, 0.f, 0.f, 0.f};
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; j++) ra[i][j] = (okp && c + j < cred) ? Asrc[base + j] : 0.f;
                    }
                }
            } else {  // FWD_NCHW / DGRAD_NCHW: lanes walk pixels
                int ih, iw, lim_h, lim_w, cred;
                if constexpr (FW) {
                    ih = a_h[0] + kx * p.vd; iw = a_w[0] + ky * p.hd;
                    lim_h = p.H; lim_w = p.W; cred = p.C;
                } else {
                    ih = a_h[0] + (x.rh + p.vp - kx * p.vd) / p.vs;
                    iw = a_w[0] + (x.rw + p.hp - ky * p.hd) / p.hs;
                    lim_h = p.Ho; lim_w = p.Wo; cred = p.KN;
                }
                const bool okp = (unsigned)ih < (unsigned)lim_h && (unsigned)iw < (unsigned)lim_w;
                const long long plane = (long long)lim_h * lim_w;
                const long long base = (long long)a_b[0] * cred * plane + (long long)ih * lim_w + iw;
#pragma unroll
                for (int i = 0; i < PA; i++) {
                    const int c = r0 + ld_kq<AK, BM, BK>(tid, i) * 4;
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        ra[i][j] = (okp && c + j < cred) ? Asrc[base + (long long)(c + j) * plane] : 0.f;
                }
            }
            // B operand: weights in [c][kh][kw][kn] order
            if constexpr (FW) {  // rows = kn, lanes walk kn
                const int n = b_row[0];
#pragma unroll
                for (int i = 0; i < PB; i++) {
                    const int c = r0 + ld_kq<BKIND, BN, BK>(tid, i) * 4;
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        rb[i][j] = (n >= 0 && c + j < p.C)
                                       ? Bsrc[((long long)(c + j) * p.taps + tap) * p.KN + n] : 0.f;
                }
            } else {  // DGRAD: rows = c, contiguous along kn (the reduction)
                const int kk = r0 + ld_kq<BKIND, BN, BK>(tid, 0) * 4;
#pragma unroll
                for (int i = 0; i < PB; i++) {
                    const int c = b_row[i];
                    const long long base = ((long long)c * p.taps + tap) * p.KN + kk;
                    if (p.vecB) {
                        rb[i] = (c >= 0 && kk < p.KN) ? *(const f32x4 *)(Bsrc + base) : f32x4{0.f, 0.f, 0.f, 0.f};
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; j++) rb[i][j] = (c >= 0 && kk + j < p.KN) ? Bsrc[base + j] : 0.f;
                    }
                }
            }
        } else {
            // WGRAD: K = output pixels p0 .. p0+BK
            const int p0 = kt * BK;
            const int P = (int)p.P;
            if constexpr (OP == CG_WGRAD_NHWC) {
                const int c = a_h[0];
#pragma unroll
                for (int i = 0; i < PA; i++) {
                    int pp = p0 + ld_kq<AK, BM, BK>(tid, i) * 4;
                    int b = 0, xo = 0, yo = 0;
                    if (pp < P) {
                        b = fdiv(pp, x.howo);
                        int pix = pp - b * HoWo;
                        xo = fdiv(pix, x.wo);
                        yo = pix - xo * p.Wo;
                    }
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int ih = xo * p.vs - p.vp + wg_kx * p.vd;
                        const int iw = yo * p.hs - p.hp + wg_ky * p.hd;
                        const bool ok = c >= 0 && pp + j < P && (unsigned)ih < (unsigned)p.H &&
                                        (unsigned)iw < (unsigned)p.W;
                        ra[i][j] = ok ? Asrc[(((long long)b * p.H + ih) * p.W + iw) * p.C + c] : 0.f;
                        if (++yo == p.Wo) { yo = 0; if (++xo == p.Ho) { xo = 0; ++b; } }
                    }
                }
                const int n = b_row[0];
#pragma unroll
                for (int i = 0; i < PB; i++) {
                    const int pp = p0 + ld_kq<BKIND, BN, BK>(tid, i) * 4;
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        rb[i][j] = (n >= 0 && pp + j < P) ? Bsrc[(long long)(pp + j) * p.KN + n] : 0.f;
                }
            } else {  // WGRAD_NCHW: both operands contiguous along pixels
                const int pp = p0 + ld_kq<AK, BM, BK>(tid, 0) * 4;
                int b = 0, xo = 0, yo = 0;
                if (pp < P) {
                    b = fdiv(pp, x.howo);
                    int pix = pp - b * HoWo;
                    xo = fdiv(pix, x.wo);
                    yo = pix - xo * p.Wo;
                }
                long long xoff[4], yoff[4];
                bool okx[4], oky[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int ih = xo * p.vs - p.vp + wg_kx * p.vd;
                    const int iw = yo * p.hs - p.hp + wg_ky * p.hd;
                    oky[j] = pp + j < P;
                    okx[j] = oky[j] && (unsigned)ih < (unsigned)p.H && (unsigned)iw < (unsigned)p.W;
                    xoff[j] = (long long)b * p.C * p.H * p.W + (long long)ih * p.W + iw;
                    yoff[j] = (long long)b * p.KN * HoWo + xo * p.Wo + yo;
                    if (++yo == p.Wo) { yo = 0; if (++xo == p.Ho) { xo = 0; ++b; } }
                }
#pragma unroll
                for (int i = 0; i < PA; i++) {
                    const int c = a_h[i];
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        ra[i][j] = (c >= 0 && okx[j]) ? Asrc[xoff[j] + (long long)c * p.H * p.W] : 0.f;
                }
#pragma unroll
                for (int i = 0; i < PB; i++) {
                    const int n = b_row[i];
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        rb[i][j] = (n >= 0 && oky[j]) ? Bsrc[yoff[j] + (long long)n * HoWo] : 0.f;
                }
            }
        }
    };

    auto store_tile = [&](int buf) {
        float *dA = sA + buf * BM * PITCH;
        float *dB = sB + buf * BN * PITCH;
#pragma unroll
        for (int i = 0; i < PA; i++)
            *(f32x4 *)(dA + ld_row<AK, BM, BK>(tid, i) * PITCH + ld_kq<AK, BM, BK>(tid, i) * 4) = ra[i];
#pragma unroll
        for (int i = 0; i < PB; i++)
            *(f32x4 *)(dB + ld_row<BKIND, BN, BK>(tid, i) * PITCH + ld_kq<BKIND, BN, BK>(tid, i) * 4) = rb[i];
    };

    // ---- MFMA main loop ----
    const int wm = wave % WM, wn = wave / WM;
    const int roff = SW ? wn * WTN : wm * WTM;  // wave origin along MFMA rows (within tile)
    const int coff = SW ? wm * WTM : wn * WTN;
    f32x16 acc[TR][TC];
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int b = 0; b < TC; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[a][b][v] = 0.f;

    auto compute_tile = [&](int buf) {
        const float *Lr = (SW ? sB + buf * BN * PITCH : sA + buf * BM * PITCH);
        const float *Lc = (SW ? sA + buf * BM * PITCH : sB + buf * BN * PITCH);
        const int half = (lane >> 5) * 8, l32 = lane & 31;
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            f32x4 ar[TR][2], ac[TC][2];
#pragma unroll
            for (int a = 0; a < TR; a++) {
                const float *s = Lr + (roff + a * 32 + l32) * PITCH + kk + half;
                ar[a][0] = *(const f32x4 *)s;
                ar[a][1] = *(const f32x4 *)(s + 4);
            }
#pragma unroll
            for (int b = 0; b < TC; b++) {
                const float *s = Lc + (coff + b * 32 + l32) * PITCH + kk + half;
                ac[b][0] = *(const f32x4 *)s;
                ac[b][1] = *(const f32x4 *)(s + 4);
            }
            // Lane half h supplies reduction index kk + h*8 + s at MFMA step s:
            // the two halves cover kk .. kk+15 over the 8 steps.
#pragma unroll
            for (int s = 0; s < 8; s++)
#pragma unroll
                for (int a = 0; a < TR; a++)
#pragma unroll
                    for (int b = 0; b < TC; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[a][s >> 2][s & 3], ac[b][s >> 2][s & 3],
                                                                        acc[a][b], 0, 0, 0);
        }
    };

    if (kt_begin < kt_end) {
        load_tile(kt_begin);
        store_tile(0);
        __syncthreads();
        for (int kt = kt_begin; kt < kt_end; ++kt) {
            const int cur = (kt - kt_begin) & 1;
            const bool has_next = kt + 1 < kt_end;
            if (has_next) load_tile(kt + 1);
            compute_tile(cur);
            if (has_next) store_tile(cur ^ 1);
            __syncthreads();
        }
    }

    // ---- write-back ----
    const int ncls = p.ncls_h * p.ncls_w;
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int b = 0; b < TC; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int ir = roff + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int ic = coff + b * 32 + (lane & 31);
                const int ml = SW ? ic : ir, nl = SW ? ir : ic;
                int m, n = nt * BN + nl;
                bool ok;
                if constexpr (WG) {
                    const int c = wg_c0 + ml;
                    m = wg_tap * p.C + c;
                    ok = c < p.C && n < p.KN;
                } else {
                    m = mt * BM + ml;
                    ok = m < x.Mrows && n < (DG ? p.C : p.KN);
                }
                if (!ok) continue;
                if (p.splits > 1)
                    p.ws[(((long long)split * ncls + cls) * p.M + m) * p.N + n] = acc[a][b][v];
                else
                    store_out<OP>(p, x, m, n, acc[a][b][v]);
            }
}

// ---- split-K reduction: one thread per (class, m, n) ----
template <int OP>
__global__ void __launch_bounds__(256) splitk_reduce_kernel(const cg_kparams p) {
    const int cls = blockIdx.y;
    const Ctx x = make_ctx(p, cls);
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)x.Mrows * p.N;
    if (idx >= total) return;
    const int m = (int)(idx / p.N), n = (int)(idx - (long long)m * p.N);
    const int ncls = p.ncls_h * p.ncls_w;
    float s = 0.f;
    for (int z = 0; z < p.splits; z++) s += p.ws[(((long long)z * ncls + cls) * p.M + m) * p.N + n];
    store_out<OP>(p, x, m, n, s);
}

__global__ void scale_kernel(float *out, long long n, float beta) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = beta == 0.0f ? 0.0f : beta * out[i];
}

// 32x32 tiled transpose through LDS: dst[c][r] = src[r][c].
__global__ void transpose_kernel(const float *__restrict__ src, float *__restrict__ dst, int R, int C) {
    __shared__ float t[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int i = threadIdx.y; i < 32; i += blockDim.y) {
        int r = r0 + i, c = c0 + threadIdx.x;
        if (r < R && c < C) t[i][threadIdx.x] = src[(long long)r * C + c];
    }
    __syncthreads();
    for (int i = threadIdx.y; i < 32; i += blockDim.y) {
        int c = c0 + i, r = r0 + threadIdx.x;
        if (r < R && c < C) dst[(long long)c * R + r] = t[threadIdx.x][i];
    }
}

// ---- configuration table ----
struct CfgEntry {
    cg_tilecfg t;
};

static const cg_tilecfg kCfgF32[] = {
    {128, 128, 16, 2, 2},  // 0
    {128, 64, 16, 2, 2},   // 1
    {64, 128, 16, 2, 2},   // 2
    {64, 64, 16, 2, 2},    // 3
    {256, 64, 16, 4, 1},   // 4
    {256, 128, 16, 2, 2},  // 5
    {64, 64, 32, 2, 2},    // 6
    {128, 64, 32, 2, 2},   // 7
    {64, 128, 32, 2, 2},   // 8
};
constexpr int kNumCfgF32 = sizeof(kCfgF32) / sizeof(kCfgF32[0]);

template <int BM, int BN, int BK, int WM, int WN, int OP>
static int launch_one(const cg_kparams *kp, hipStream_t st) {
    dim3 grid(kp->mtiles * kp->ntiles, kp->ncls_h * kp->ncls_w, kp->splits);
    hipLaunchKernelGGL((igemm_f32_kernel<BM, BN, BK, WM, WN, OP>), grid, dim3(NT), 0, st, *kp);
    return (int)hipGetLastError();
}

template <int BM, int BN, int BK, int WM, int WN>
static int launch_cfg(const cg_kparams *kp, hipStream_t st) {
    switch (kp->op) {
    case CG_FWD_NHWC: return launch_one<BM, BN, BK, WM, WN, CG_FWD_NHWC>(kp, st);
    case CG_FWD_NCHW: return launch_one<BM, BN, BK, WM, WN, CG_FWD_NCHW>(kp, st);
    case CG_WGRAD_NHWC: return launch_one<BM, BN, BK, WM, WN, CG_WGRAD_NHWC>(kp, st);
    case CG_WGRAD_NCHW: return launch_one<BM, BN, BK, WM, WN, CG_WGRAD_NCHW>(kp, st);
    case CG_DGRAD_NHWC: return launch_one<BM, BN, BK, WM, WN, CG_DGRAD_NHWC>(kp, st);
    case CG_DGRAD_NCHW: return launch_one<BM, BN, BK, WM, WN, CG_DGRAD_NCHW>(kp, st);
    }
    return (int)hipErrorInvalidValue;
}

}  // namespace cg

using namespace cg;

extern "C" int cg_num_tilecfgs(int elem) { return elem == CG_ELEM_F32 ? kNumCfgF32 : 0; }

extern "C" int cg_get_tilecfg(int elem, int cfg, cg_tilecfg *out) {
    if (elem != CG_ELEM_F32 || cfg < 0 || cfg >= kNumCfgF32) return -1;
    *out = kCfgF32[cfg];
    return 0;
}

extern "C" int cg_tilecfg_supports(int elem, int cfg, int op) {
    if (elem != CG_ELEM_F32 || cfg < 0 || cfg >= kNumCfgF32 || op < 0 || op >= CG_NUM_OPKINDS) return 0;
    return 1;  // every F32 configuration instantiates every operation
}

extern "C" int cg_launch_igemm(int elem, int cfg, const cg_kparams *kp, void *stream) {
    if (elem != CG_ELEM_F32) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    switch (cfg) {
    case 0: return launch_cfg<128, 128, 16, 2, 2>(kp, st);
    case 1: return launch_cfg<128, 64, 16, 2, 2>(kp, st);
    case 2: return launch_cfg<64, 128, 16, 2, 2>(kp, st);
    case 3: return launch_cfg<64, 64, 16, 2, 2>(kp, st);
    case 4: return launch_cfg<256, 64, 16, 4, 1>(kp, st);
    case 5: return launch_cfg<256, 128, 16, 2, 2>(kp, st);
    case 6: return launch_cfg<64, 64, 32, 2, 2>(kp, st);
    case 7: return launch_cfg<128, 64, 32, 2, 2>(kp, st);
    case 8: return launch_cfg<64, 128, 32, 2, 2>(kp, st);
    }
    return (int)hipErrorInvalidValue;
}

extern "C" int cg_launch_splitk_reduce(const cg_kparams *kp, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    // rows of the largest class bound every class
    long long total = (long long)kp->M * kp->N;
    dim3 grid((unsigned)((total + 255) / 256), kp->ncls_h * kp->ncls_w);
    switch (kp->op) {
    case CG_FWD_NHWC: hipLaunchKernelGGL(splitk_reduce_kernel<CG_FWD_NHWC>, grid, dim3(256), 0, st, *kp); break;
    case CG_FWD_NCHW: hipLaunchKernelGGL(splitk_reduce_kernel<CG_FWD_NCHW>, grid, dim3(256), 0, st, *kp); break;
    case CG_WGRAD_NHWC: hipLaunchKernelGGL(splitk_reduce_kernel<CG_WGRAD_NHWC>, grid, dim3(256), 0, st, *kp); break;
    case CG_WGRAD_NCHW: hipLaunchKernelGGL(splitk_reduce_kernel<CG_WGRAD_NCHW>, grid, dim3(256), 0, st, *kp); break;
    case CG_DGRAD_NHWC: hipLaunchKernelGGL(splitk_reduce_kernel<CG_DGRAD_NHWC>, grid, dim3(256), 0, st, *kp); break;
    case CG_DGRAD_NCHW: hipLaunchKernelGGL(splitk_reduce_kernel<CG_DGRAD_NCHW>, grid, dim3(256), 0, st, *kp); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

extern "C" int cg_launch_scale(float *out, long long n, float beta, void *stream) {
    if (n <= 0) return 0;
    long long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, out, n, beta);
    return (int)hipGetLastError();
}

extern "C" int cg_launch_transpose_f32(const float *src, float *dst, int R, int C, void *stream) {
    if (R <= 0 || C <= 0) return 0;
    dim3 grid((C + 31) / 32, (R + 31) / 32);
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, (hipStream_t)stream, src, dst, R, C);
    return (int)hipGetLastError();
}
