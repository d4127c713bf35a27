// igemm_kernel.hpp -- the implicit-GEMM convolution kernel template for gfx950.
//
// One template serves the six operations of convGemm (src/convGemm.c:48-146)
// and three operand types (FP32, BF16, FP8-e4m3; FP32 accumulation).  The
// reference packs im2row/im2col panels on the CPU (src/im2row_nhwc.c:28-145,
// src/im2col_nchw.c:29-144, :238-335) and runs a BLIS micro-kernel on them;
// here each workgroup gathers its convolution patches straight from HBM
// (buffer loads: out-of-range taps and padding read as 0 in hardware) into an
// LDS double buffer and contracts them on MFMA.  The reference's post-process
// callbacks (bias/BN/ReLU epilogue, NCHW transposed store, row2im/col2im
// scatter) become the write-back of the accumulators: backward data is a
// gather GEMM per stride residue class (one writer per dX element, no
// atomics, no multiply by a structural zero).
//
// Tile: BM x BN outputs per 256-thread workgroup, BK reduction elements per
// K step, 4 waves arranged WM x WN, each wave owning (BM/WM) x (BN/WN) as
// 32x32 MFMA blocks.  LDS rows are K-contiguous, padded by one 16-byte chunk.
#pragma once
#include <type_traits>

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
// Forward weights: every type reads a [kn][kh][kw][c] repack, so that a
// 16-byte chunk of 4/8/16 channels is one vector load.  Returning false makes
// the loader read the caller's [c][kh][kw][kn] order in place (lanes walk kn).
template <typename T> constexpr bool fwd_b_kmajor() { return true; }

// How each operand tile is read from HBM:
//   KV: contiguous along K, one 16-byte chunk per load (VEC), else element loads
//   KS: contiguous along K, element loads (unaligned runs)
//   RS: contiguous along the row dimension (lanes walk rows), element loads
template <int OP> constexpr int a_kind() {
    return OP == CG_FWD_NHWC ? LK_KV : OP == CG_FWD_NCHW ? LK_RS : OP == CG_DGRAD_NHWC ? LK_KV
           : OP == CG_DGRAD_NCHW ? LK_RS : OP == CG_WGRAD_NHWC ? LK_RS : LK_KS;
}
template <typename T, int OP> constexpr int b_kind() {
    return is_fwd<OP>() ? (fwd_b_kmajor<T>() ? LK_KV : LK_RS) : is_dgrad<OP>() ? LK_KV
           : OP == CG_WGRAD_NHWC ? LK_RS : LK_KS;
}

template <int KIND, int R, int KQ> constexpr bool loader_ok() {
    if (R * KQ < NT) return false;
    if (KIND == LK_RS) return NT % R == 0;
    return (R * KQ) % NT == 0;
}

// Weight gradient, NHWC, 16-byte operands ("RV" loader): both operands are
// contiguous along their row dimension (channels of X, kn of dY) and the K
// dimension is pixels.  A thread loads one 16-byte group of VW rows at J
// consecutive pixels (J = (R/VW)*BK/NT vector loads) and writes it transposed:
// VW rows, each a K-contiguous run of J elements.  Consecutive lanes read
// consecutive 16-byte groups of one pixel (coalesced).  In LDS the 16-byte
// chunk index of a row is XOR-swizzled by (row >> 3), which spreads the
// transposed stores over the banks; 8 consecutive rows share a swizzle, so the
// MFMA operand reads stay conflict-free.
template <int R, int BK, int VW> constexpr bool rv_ok() {
    if (R % VW != 0 || R / VW == 0 || NT % (R / VW) != 0 || (R / VW) * BK % NT != 0) return false;
    const int J = (R / VW) * BK / NT;
    return J >= 1 && (J <= VW ? VW % J == 0 : J % VW == 0);
}

template <int KQ, int VW>
__device__ inline int rv_k(int row, int k) {  // LDS element offset of K index k within `row`
    return ((k / VW) ^ ((row >> 3) & (KQ - 1))) * VW + (k % VW);
}

template <int NB>
__device__ inline void lds_put(void *dst, const void *src) {
    if constexpr (NB == 16) {
        i32x4 w;
        __builtin_memcpy(&w, src, 16);
        *(i32x4 *)dst = w;
    } else if constexpr (NB == 8) {
        i32x2 w;
        __builtin_memcpy(&w, src, 8);
        *(i32x2 *)dst = w;
    } else if constexpr (NB == 4) {
        int w;
        __builtin_memcpy(&w, src, 4);
        *(int *)dst = w;
    } else {
        static_assert(NB == 2, "piece size");
        unsigned short w;
        __builtin_memcpy(&w, src, 2);
        *(unsigned short *)dst = w;
    }
}

// Element r of each of the J chunks in `src` forms row (row0 + r)'s K run
// [k0, k0 + J), written in pieces of at most one 16-byte chunk.
template <typename T, int J, int VW, int PITCH, int KQ, typename V>
__device__ __attribute__((always_inline)) inline void rv_store(T *dst, int row0, int k0, const V (&src)[J]) {
    constexpr int PIECE = J < VW ? J : VW;
#pragma unroll
    for (int r = 0; r < VW; r++) {
        const int row = row0 + r;
#pragma unroll
        for (int h = 0; h < J; h += PIECE) {
            T v[PIECE];
#pragma unroll
            for (int i = 0; i < PIECE; i++) v[i] = src[h + i][r];
            lds_put<PIECE * (int)sizeof(T)>(dst + row * PITCH + rv_k<KQ, VW>(row, k0 + h), v);
        }
    }
}

// Thread -> (row, chunk) of a loader.
template <int KIND, int R, int KQ>
__device__ inline int ld_row(int tid, int i) {
    if constexpr (KIND == LK_RS) return tid % R;
    else return tid / KQ + i * (NT / KQ);
}
template <int KIND, int R, int KQ>
__device__ inline int ld_kq(int tid, int i) {
    if constexpr (KIND == LK_RS) return tid / R + i * (NT / R);
    else return tid % KQ;
}

constexpr int FAR = -(1 << 29);  // row coordinate that fails every bounds test

// One workgroup item of problem p: an output tile, split-K slice and (DGRAD)
// stride class; `lin` is the item's linear index within the problem.
// Stream-K (sk_k0 >= 0; FWD/WGRAD): `lin` is the tile, K tiles
// [sk_k0, sk_k1) are reduced, and the accumulators go to `piece` ([BM][BN],
// tile-local) unless piece is null (the whole K range: direct write-back).
template <typename T, int BM, int BN, int BK, int WM, int WN, int OP, bool VEC>
__device__ __attribute__((always_inline)) inline void igemm_tile(const cg_kparams &p, const int lin, const int sk_k0 = -1,
                                                                 const int sk_k1 = 0, float *piece = nullptr) {
    typedef typename ET<T>::V V;
    constexpr int VW = ET<T>::VEC;
    constexpr int G = ET<T>::G;
    constexpr int ES = sizeof(T);
    constexpr int PITCH = BK + VW;
    constexpr int KQ = BK / VW;
    constexpr int PA = BM * KQ / NT;
    constexpr int PB = BN * KQ / NT;
    constexpr int AK = a_kind<OP>();
    constexpr int BKIND = b_kind<T, OP>();
    constexpr bool FW = is_fwd<OP>(), DG = is_dgrad<OP>(), WG = is_wgrad<OP>();
    constexpr bool SW = swap_out<OP>();
    constexpr int NRA = (AK == LK_RS) ? 1 : PA;  // distinct A rows per thread
    constexpr int NRB = (BKIND == LK_RS) ? 1 : PB;
    static_assert(loader_ok<AK, BM, KQ>() && loader_ok<BKIND, BN, KQ>(), "loader geometry");
    // RV loader (weight gradient, NHWC, 16-byte operands); its per-thread chunk
    // counts are PA and PB, the same as the element loader's
    constexpr bool RV = OP == CG_WGRAD_NHWC && VEC && rv_ok<BM, BK, VW>() && rv_ok<BN, BK, VW>();
    static_assert(!RV || (KQ & (KQ - 1)) == 0, "RV swizzle needs a power-of-two chunk count");
    static_assert(BK % G == 0, "BK must be a multiple of the MFMA group");
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int ER = SW ? WTN : WTM, EC = SW ? WTM : WTN;
    constexpr int TR = ER / 32, TC = EC / 32;
    static_assert(TR >= 1 && TC >= 1 && ER % 32 == 0 && EC % 32 == 0, "wave tile");

    __shared__ __attribute__((aligned(16))) T sA[2 * BM * PITCH];
    __shared__ __attribute__((aligned(16))) T sB[2 * BN * PITCH];
    __shared__ int s_kx[64], s_ky[64], s_nk[2];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // ---- item -> (class, tile, split).  The N tiles of one M panel are
    // adjacent: the gathered operand (the large one) is then fetched from HBM
    // once and the small weight panel stays L2-resident (with M fastest the
    // gathered operand was measured re-read once per N tile).  Backward data:
    // the residue classes of one tile are adjacent items, so every XCD's run
    // of items (xcd_run) holds all classes in equal shares -- with the class
    // outermost, each XCD ran one class and the XCDs with the most taps per
    // class set the time -- and within a tile the class with the most taps
    // comes first (cls_order).
    const int tiles = p.mtiles * p.ntiles;
    const bool skm = sk_k0 >= 0;
    const int ncls_all = p.ncls_h * p.ncls_w;
    const int cslot = skm ? 0 : lin % ncls_all, rest0 = skm ? 0 : lin / ncls_all;
    const int tile = skm ? lin : rest0 % tiles, rest = skm ? 0 : rest0 / tiles;
    const int split = rest % p.splits;
    const int cls = p.cls_order ? (int)((p.cls_order >> (4 * cslot)) & 15) : cslot;
    int mt, nt;
    tile_mn(p, tile, mt, nt);

    const Ctx x = make_ctx(p, cls);
    if (!WG && mt * BM >= x.Mrows) return;  // uniform: DGRAD classes smaller than the grid

    const __amdgpu_buffer_rsrc_t rA = make_rsrc(p.A, p.a_bytes);
    const __amdgpu_buffer_rsrc_t rB = make_rsrc(p.Bm, p.b_bytes);
    const int HoWo = p.Ho * p.Wo;

    // ---- reduction range of this block ----
    int nky = p.KW, nkx = p.KH;
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
    const int per_split = (KT + p.splits - 1) / p.splits;
    const int kt_begin = skm ? sk_k0 : min(KT, split * per_split);
    const int kt_end = skm ? sk_k1 : min(KT, kt_begin + per_split);

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
        const int row = RV ? (tid % (BM / VW)) * VW : ld_row<AK, BM, KQ>(tid, i);
        a_b[i] = 0;
        a_h[i] = FAR;
        a_w[i] = FAR;
        if constexpr (FW) {
            const int m = mt * BM + row;
            if (m < (int)p.P) {
                const int b = fdiv(m, x.howo);
                const int pix = m - b * HoWo;
                const int xo = fdiv(pix, x.wo);
                a_b[i] = b;
                a_h[i] = xo * p.vs - p.vp;
                a_w[i] = (pix - xo * p.Wo) * p.hs - p.hp;
            }
        } else if constexpr (DG) {
            const int q = mt * BM + row;
            if (q < x.Mrows) {
                const int b = fdiv(q, x.cls_hw);
                const int rem2 = q - b * (x.Hc * x.Wc);
                const int ii = fdiv(rem2, x.cls_w);
                a_b[i] = b;
                a_h[i] = ii;
                a_w[i] = rem2 - ii * x.Wc;
            }
        } else {
            const int c = wg_c0 + row;
            a_h[i] = c < p.C ? c : -1;
        }
    }
    int b_row[NRB];  // FWD/WGRAD: kn index (or -1); DGRAD: c index (or -1)
#pragma unroll
    for (int i = 0; i < NRB; i++) {
        const int r = nt * BN + (RV ? (tid % (BN / VW)) * VW : ld_row<BKIND, BN, KQ>(tid, i));
        b_row[i] = r < (DG ? p.C : p.KN) ? r : -1;
    }

    // Two register stages: the loads of K tile t+2 are in flight while tile t
    // is multiplied out of LDS and tile t+1 is written to the other LDS buffer.
    V ra0[PA], rb0[PB], ra1[PA], rb1[PB];

    // reduction iterator (FWD/DGRAD): tile kt -> (tap index ti, channel chunk ch)
    int it_ch = 0, it_ti = 0, it_kx = 0, it_ky = 0, it_vx = 0, it_vy = 0;
    auto it_set = [&](int kt) __attribute__((always_inline)) {
        it_ti = kt / p.kchunks;
        it_ch = kt - it_ti * p.kchunks;
        if constexpr (DG) {
            if (nky > 0) {
                it_vx = it_ti / nky;
                it_vy = it_ti - it_vx * nky;
                it_kx = s_kx[it_vx];
                it_ky = s_ky[it_vy];
            }
        } else {
            it_kx = it_ti / p.KW;
            it_ky = it_ti - it_kx * p.KW;
        }
    };
    auto it_next = [&]() __attribute__((always_inline)) {
        if (++it_ch == p.kchunks) {
            it_ch = 0;
            ++it_ti;
            if constexpr (DG) {
                if (++it_vy == nky) { it_vy = 0; ++it_vx; }
                it_kx = s_kx[min(it_vx, 63)];
                it_ky = s_ky[it_vy];
            } else {
                if (++it_ky == p.KW) { it_ky = 0; ++it_kx; }
            }
        }
    };

    // ---- gather one K tile (global -> registers) ----
    auto load_tile = [&](int kt, V (&ra)[PA], V (&rb)[PB]) __attribute__((always_inline)) {
        if constexpr (!WG) {
            const int r0 = it_ch * BK;
            const int kx = it_kx, ky = it_ky;
            const int tap = kx * p.KW + ky;
            const int cred = FW ? p.C : p.KN;  // reduction channels
            int dh, dw, lim_h, lim_w;
            if constexpr (FW) {
                dh = kx * p.vd; dw = ky * p.hd; lim_h = p.H; lim_w = p.W;
            } else {
                dh = (x.rh + p.vp - kx * p.vd) / p.vs;
                dw = (x.rw + p.hp - ky * p.hd) / p.hs;
                lim_h = p.Ho; lim_w = p.Wo;
            }
            // A operand
            if constexpr (AK == LK_KV) {  // NHWC: chunk = VW consecutive channels of one pixel
                const int c = r0 + ld_kq<AK, BM, KQ>(tid, 0) * VW;
#pragma unroll
                for (int i = 0; i < PA; i++) {
                    const int ih = a_h[i] + dh, iw = a_w[i] + dw;
                    const bool okp = (unsigned)ih < (unsigned)lim_h && (unsigned)iw < (unsigned)lim_w;
                    const unsigned base =
                        (((unsigned)a_b[i] * lim_h + (unsigned)ih) * lim_w + (unsigned)iw) * cred + c;
                    if constexpr (VEC) {
                        ra[i] = ld_vec<T>(rA, (okp && c < cred) ? base * ES : OOB);
                    } else {
#pragma unroll
                        for (int j = 0; j < VW; j++)
                            ra[i][j] = ld_one<T>(rA, (okp && c + j < cred) ? (base + j) * ES : OOB);
                    }
                }
            } else {  // NCHW: lanes walk pixels, one plane per channel
                const int ih = a_h[0] + dh, iw = a_w[0] + dw;
                const bool okp = (unsigned)ih < (unsigned)lim_h && (unsigned)iw < (unsigned)lim_w;
                const unsigned plane = (unsigned)lim_h * lim_w * ES;  // bytes between channels
                const unsigned base =
                    ((unsigned)a_b[0] * cred * lim_h * lim_w + (unsigned)ih * lim_w + (unsigned)iw) * ES;
#pragma unroll
                for (int i = 0; i < PA; i++) {
                    const int c = r0 + ld_kq<AK, BM, KQ>(tid, i) * VW;
                    const unsigned cb = base + (unsigned)c * plane;
#pragma unroll
                    for (int j = 0; j < VW; j++)
                        ra[i][j] = ld_one<T>(rA, (okp && c + j < cred) ? cb + j * plane : OOB);
                }
            }
            // B operand
            if constexpr (FW && !fwd_b_kmajor<T>()) {  // [c][tap][kn]: lanes walk kn
                const int n = b_row[0];
                const unsigned cstride = (unsigned)p.taps * p.KN * ES;  // bytes between channels
                const unsigned nbase = ((unsigned)tap * p.KN + (unsigned)n) * ES;
#pragma unroll
                for (int i = 0; i < PB; i++) {
                    const int c = r0 + ld_kq<BKIND, BN, KQ>(tid, i) * VW;
                    const unsigned base = (unsigned)c * cstride + nbase;
#pragma unroll
                    for (int j = 0; j < VW; j++)
                        rb[i][j] = ld_one<T>(rB, (n >= 0 && c + j < p.C) ? base + j * cstride : OOB);
                }
            } else {  // FWD LP: [kn][tap][c];  DGRAD: [c][tap][kn] -- contiguous along the reduction
                const int k0 = r0 + ld_kq<BKIND, BN, KQ>(tid, 0) * VW;
#pragma unroll
                for (int i = 0; i < PB; i++) {
                    const int row = b_row[i];
                    const unsigned base = ((unsigned)row * p.taps + tap) * cred + k0;
                    if constexpr (VEC) {
                        rb[i] = ld_vec<T>(rB, (row >= 0 && k0 < cred) ? base * ES : OOB);
                    } else {
#pragma unroll
                        for (int j = 0; j < VW; j++)
                            rb[i][j] = ld_one<T>(rB, (row >= 0 && k0 + j < cred) ? (base + j) * ES : OOB);
                    }
                }
            }
        } else {
            // WGRAD: K = output pixels p0 .. p0+BK
            const int p0 = kt * BK;
            const int P = (int)p.P;
            if constexpr (RV) {
                const int c = a_h[0];  // first channel of this thread's 16-byte group (or -1)
                const int pa = p0 + (tid / (BM / VW)) * PA;
                int b = 0, xo = 0, yo = 0;
                if (pa < P) {
                    b = fdiv(pa, x.howo);
                    const int pix = pa - b * HoWo;
                    xo = fdiv(pix, x.wo);
                    yo = pix - xo * p.Wo;
                }
#pragma unroll
                for (int i = 0; i < PA; i++) {
                    const int ih = xo * p.vs - p.vp + wg_kx * p.vd;
                    const int iw = yo * p.hs - p.hp + wg_ky * p.hd;
                    const bool ok = c >= 0 && pa + i < P && (unsigned)ih < (unsigned)p.H && (unsigned)iw < (unsigned)p.W;
                    ra[i] = ld_vec<T>(rA, ok ? ((((unsigned)b * p.H + ih) * p.W + iw) * p.C + c) * ES : OOB);
                    if (++yo == p.Wo) { yo = 0; if (++xo == p.Ho) { xo = 0; ++b; } }
                }
                const int n = b_row[0];
                const int pb = p0 + (tid / (BN / VW)) * PB;
#pragma unroll
                for (int i = 0; i < PB; i++)
                    rb[i] = ld_vec<T>(rB, (n >= 0 && pb + i < P) ? ((unsigned)(pb + i) * p.KN + n) * ES : OOB);
            } else if constexpr (OP == CG_WGRAD_NHWC) {
                const int c = a_h[0];
#pragma unroll
                for (int i = 0; i < PA; i++) {
                    const int pp = p0 + ld_kq<AK, BM, KQ>(tid, i) * VW;
                    int b = 0, xo = 0, yo = 0;
                    if (pp < P) {
                        b = fdiv(pp, x.howo);
                        const int pix = pp - b * HoWo;
                        xo = fdiv(pix, x.wo);
                        yo = pix - xo * p.Wo;
                    }
#pragma unroll
                    for (int j = 0; j < VW; j++) {
                        const int ih = xo * p.vs - p.vp + wg_kx * p.vd;
                        const int iw = yo * p.hs - p.hp + wg_ky * p.hd;
                        const bool ok = c >= 0 && pp + j < P && (unsigned)ih < (unsigned)p.H &&
                                        (unsigned)iw < (unsigned)p.W;
                        ra[i][j] = ld_one<T>(rA, ok ? ((((unsigned)b * p.H + ih) * p.W + iw) * p.C + c) * ES : OOB);
                        if (++yo == p.Wo) { yo = 0; if (++xo == p.Ho) { xo = 0; ++b; } }
                    }
                }
                const int n = b_row[0];
#pragma unroll
                for (int i = 0; i < PB; i++) {
                    const int pp = p0 + ld_kq<BKIND, BN, KQ>(tid, i) * VW;
#pragma unroll
                    for (int j = 0; j < VW; j++)
                        rb[i][j] = ld_one<T>(rB, (n >= 0 && pp + j < P) ? ((unsigned)(pp + j) * p.KN + n) * ES : OOB);
                }
            } else {  // WGRAD_NCHW: both operands contiguous along pixels
                const int pp = p0 + ld_kq<AK, BM, KQ>(tid, 0) * VW;
                int b = 0, xo = 0, yo = 0;
                if (pp < P) {
                    b = fdiv(pp, x.howo);
                    const int pix = pp - b * HoWo;
                    xo = fdiv(pix, x.wo);
                    yo = pix - xo * p.Wo;
                }
                unsigned xoff[VW], yoff[VW];
                bool okx[VW], oky[VW];
#pragma unroll
                for (int j = 0; j < VW; j++) {
                    const int ih = xo * p.vs - p.vp + wg_kx * p.vd;
                    const int iw = yo * p.hs - p.hp + wg_ky * p.hd;
                    oky[j] = pp + j < P;
                    okx[j] = oky[j] && (unsigned)ih < (unsigned)p.H && (unsigned)iw < (unsigned)p.W;
                    xoff[j] = (unsigned)b * p.C * p.H * p.W + (unsigned)ih * p.W + (unsigned)iw;
                    yoff[j] = (unsigned)b * p.KN * HoWo + (unsigned)(xo * p.Wo + yo);
                    if (++yo == p.Wo) { yo = 0; if (++xo == p.Ho) { xo = 0; ++b; } }
                }
#pragma unroll
                for (int i = 0; i < PA; i++) {
                    const int c = a_h[i];
#pragma unroll
                    for (int j = 0; j < VW; j++)
                        ra[i][j] = ld_one<T>(rA, (c >= 0 && okx[j]) ? (xoff[j] + (unsigned)c * p.H * p.W) * ES : OOB);
                }
#pragma unroll
                for (int i = 0; i < PB; i++) {
                    const int n = b_row[i];
#pragma unroll
                    for (int j = 0; j < VW; j++)
                        rb[i][j] = ld_one<T>(rB, (n >= 0 && oky[j]) ? (yoff[j] + (unsigned)n * HoWo) * ES : OOB);
                }
            }
        }
    };

    auto store_tile = [&](int buf, const V (&ra)[PA], const V (&rb)[PB]) __attribute__((always_inline)) {
        T *dA = sA + buf * BM * PITCH;
        T *dB = sB + buf * BN * PITCH;
        if constexpr (RV) {
            rv_store<T, PA, VW, PITCH, KQ>(dA, (tid % (BM / VW)) * VW, (tid / (BM / VW)) * PA, ra);
            rv_store<T, PB, VW, PITCH, KQ>(dB, (tid % (BN / VW)) * VW, (tid / (BN / VW)) * PB, rb);
            return;
        }
#pragma unroll
        for (int i = 0; i < PA; i++)
            *(V *)(dA + ld_row<AK, BM, KQ>(tid, i) * PITCH + ld_kq<AK, BM, KQ>(tid, i) * VW) = ra[i];
#pragma unroll
        for (int i = 0; i < PB; i++)
            *(V *)(dB + ld_row<BKIND, BN, KQ>(tid, i) * PITCH + ld_kq<BKIND, BN, KQ>(tid, i) * VW) = rb[i];
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

    // LDS element offset of K index k (a multiple of 4 elements) within row r
    auto kof = [](int r, int k) __attribute__((always_inline)) { return RV ? rv_k<KQ, VW>(r, k) : k; };

    auto compute_tile = [&](int buf) __attribute__((always_inline)) {
        const T *Lr = SW ? sB + buf * BN * PITCH : sA + buf * BM * PITCH;
        const T *Lc = SW ? sA + buf * BM * PITCH : sB + buf * BN * PITCH;
        const int l32 = lane & 31;
        // Lane half h supplies the reduction elements [g + h*G/2, g + (h+1)*G/2)
        // of group g; A and B use the same assignment, so any order within the
        // group gives the same sum.
        const int hoff = (lane >> 5) * (G / 2);
#pragma unroll
        for (int g = 0; g < BK; g += G) {
            if constexpr (std::is_same<T, float>::value) {
                f32x4 ar[TR][2], ac[TC][2];
#pragma unroll
                for (int a = 0; a < TR; a++) {
                    const int r = roff + a * 32 + l32;
                    ar[a][0] = *(const f32x4 *)(Lr + r * PITCH + kof(r, g + hoff));
                    ar[a][1] = *(const f32x4 *)(Lr + r * PITCH + kof(r, g + hoff + 4));
                }
#pragma unroll
                for (int b = 0; b < TC; b++) {
                    const int r = coff + b * 32 + l32;
                    ac[b][0] = *(const f32x4 *)(Lc + r * PITCH + kof(r, g + hoff));
                    ac[b][1] = *(const f32x4 *)(Lc + r * PITCH + kof(r, g + hoff + 4));
                }
#pragma unroll
                for (int s = 0; s < 8; s++)
#pragma unroll
                    for (int a = 0; a < TR; a++)
#pragma unroll
                        for (int b = 0; b < TC; b++)
                            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[a][s >> 2][s & 3],
                                                                            ac[b][s >> 2][s & 3], acc[a][b], 0, 0, 0);
            } else if constexpr (std::is_same<T, uint16_t>::value) {
                u16x8 ar[TR], ac[TC];
#pragma unroll
                for (int a = 0; a < TR; a++) {
                    const int r = roff + a * 32 + l32;
                    ar[a] = *(const u16x8 *)(Lr + r * PITCH + kof(r, g + hoff));
                }
#pragma unroll
                for (int b = 0; b < TC; b++) {
                    const int r = coff + b * 32 + l32;
                    ac[b] = *(const u16x8 *)(Lc + r * PITCH + kof(r, g + hoff));
                }
#pragma unroll
                for (int a = 0; a < TR; a++)
#pragma unroll
                    for (int b = 0; b < TC; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, ar[a]), __builtin_bit_cast(bf16x8, ac[b]), acc[a][b], 0, 0, 0);
            } else {
                i32x8 ar[TR], ac[TC];
#pragma unroll
                for (int a = 0; a < TR; a++) {
                    const uint8_t *s = Lr + (roff + a * 32 + l32) * PITCH + g + hoff;
                    const i32x4 lo = *(const i32x4 *)s, hi = *(const i32x4 *)(s + 16);
                    ar[a] = i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                }
#pragma unroll
                for (int b = 0; b < TC; b++) {
                    const uint8_t *s = Lc + (coff + b * 32 + l32) * PITCH + g + hoff;
                    const i32x4 lo = *(const i32x4 *)s, hi = *(const i32x4 *)(s + 16);
                    ac[b] = i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                }
                // operand formats 0/0 = fp8 e4m3 for A and B, E8M0 block scales 127 (= 1.0);
                // the per-tensor scales are applied in the epilogue
#pragma unroll
                for (int a = 0; a < TR; a++)
#pragma unroll
                    for (int b = 0; b < TC; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(ar[a], ac[b], acc[a][b], 0, 0, 0,
                                                                                    127, 0, 127);
            }
        }
    };

    // Tile i (counted from kt_begin) lives in register stage i&1 and LDS buffer
    // i&1.  At step i: issue the loads of tile i+2 into stage i&1 (stored to
    // LDS at step i-1), multiply LDS buffer i&1, write stage (i+1)&1 into LDS
    // buffer (i+1)&1 (last read at step i-1, before the barrier), barrier.
    auto step = [&](auto stage, int kt) __attribute__((always_inline)) {
        constexpr int S = decltype(stage)::value;
        if (kt + 2 < kt_end) {
            if constexpr (!WG) it_next();
            if constexpr (S == 0) load_tile(kt + 2, ra0, rb0);
            else load_tile(kt + 2, ra1, rb1);
        }
        compute_tile(S);
        if (kt + 1 < kt_end) {
            if constexpr (S == 0) store_tile(1, ra1, rb1);
            else store_tile(0, ra0, rb0);
        }
        __syncthreads();
    };
    if (kt_begin < kt_end) {
        if constexpr (!WG) it_set(kt_begin);
        load_tile(kt_begin, ra0, rb0);
        if (kt_begin + 1 < kt_end) {
            if constexpr (!WG) it_next();
            load_tile(kt_begin + 1, ra1, rb1);
        }
        store_tile(0, ra0, rb0);
        __syncthreads();
        for (int kt = kt_begin; kt < kt_end; kt += 2) {
            step(std::integral_constant<int, 0>{}, kt);
            if (kt + 1 < kt_end) step(std::integral_constant<int, 1>{}, kt + 1);
        }
    }

    // ---- write-back ----
    const int ncls = p.ncls_h * p.ncls_w;
    if (piece || p.splits > 1) {  // raw partials: stream-K piece (tile-local) or split-K slice
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int b = 0; b < TC; b++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int ir = roff + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                    const int ic = coff + b * 32 + (lane & 31);
                    const int ml = SW ? ic : ir, nl = SW ? ir : ic;
                    if (piece) {  // [BM][BN]; pixel-major [BN][BM] for the NCHW forward (as its output)
                        if constexpr (OP == CG_FWD_NCHW) piece[nl * BM + ml] = acc[a][b][v];
                        else piece[ml * BN + nl] = acc[a][b][v];
                        continue;
                    }
                    int m;
                    const int n = nt * BN + nl;
                    bool ok;
                    if constexpr (WG) {
                        const int c = wg_c0 + ml;
                        m = wg_tap * p.C + c;
                        ok = c < p.C && n < p.KN;
                    } else {
                        m = mt * BM + ml;
                        ok = m < x.Mrows && n < (DG ? p.C : p.KN);
                    }
                    if (ok) p.ws[(((long long)split * ncls + cls) * p.M + m) * p.N + n] = acc[a][b][v];
                }
        return;
    }
    if constexpr (!FW) {  // weight gradient, backward data
        auto pos = [&](int a, int b, int v, int &m, int &n) __attribute__((always_inline)) {
            const int ir = roff + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            const int ic = coff + b * 32 + (lane & 31);
            const int ml = SW ? ic : ir, nl = SW ? ir : ic;
            n = nt * BN + nl;
            if constexpr (WG) {
                const int c = wg_c0 + ml;
                m = wg_tap * p.C + c;
                return c < p.C && n < p.KN;
            } else {
                m = mt * BM + ml;
                return m < x.Mrows && n < p.C;
            }
        };
        if (p.beta == 0.0f) {  // nothing to load: one store per output
#pragma unroll
            for (int a = 0; a < TR; a++)
#pragma unroll
                for (int b = 0; b < TC; b++)
#pragma unroll
                    for (int v = 0; v < 16; v++) {
                        int m, n;
                        if (pos(a, b, v, m, n)) store_out<OP>(p, x, m, n, acc[a][b][v]);
                    }
            return;
        }
        // beta != 0 (backward data accumulates): old values of 4 outputs, then their
        // stores -- one wait per 4 outputs instead of one per output
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int b = 0; b < TC; b++)
                store_block<OP>(p, x, acc[a][b], nullptr, 0, [&](int v, int &m, int &n, int &nl) __attribute__((always_inline)) {
                    nl = 0;
                    return pos(a, b, v, m, n);
                });
        return;
    }
    if (!epi_staged<OP>(p) && p.beta == 0.0f) {  // plain forward: nothing to load, one store per output
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int b = 0; b < TC; b++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int ir = roff + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                    const int ic = coff + b * 32 + (lane & 31);
                    const int ml = SW ? ic : ir, nl = SW ? ir : ic;
                    const int m = mt * BM + ml, n = nt * BN + nl;
                    if (m < x.Mrows && n < p.KN) store_out<OP, (sizeof(T) < 4)>(p, x, m, n, acc[a][b][v]);
                }
        return;
    }
    // forward: the fused epilogue, per-channel vectors staged in LDS (the operand tiles are dead)
    float *s_epi = nullptr;
    if (epi_staged<OP>(p)) {
        static_assert(5 * BN * sizeof(float) <= sizeof(sA), "epilogue vectors fit the A tile's LDS");
        __syncthreads();
        s_epi = reinterpret_cast<float *>(sA);
        epi_stage(p, nt * BN, BN, s_epi, tid, NT);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int b = 0; b < TC; b++)
            store_block<OP, (sizeof(T) < 4)>(p, x, acc[a][b], s_epi, BN, [&](int v, int &m, int &n, int &nl) __attribute__((always_inline)) {
                const int ir = roff + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int ic = coff + b * 32 + (lane & 31);
                const int ml = SW ? ic : ir;
                nl = SW ? ir : ic;
                n = nt * BN + nl;
                m = mt * BM + ml;
                return m < x.Mrows && n < p.KN;
            });
    if (s_epi) __syncthreads();  // the LDS is rewritten by the next item of a persistent launch
}


template <typename T, int BM, int BN, int BK, int WM, int WN, int OP, bool VEC>
__global__ void __launch_bounds__(NT) igemm_kernel(const cg_kparams p) {
    igemm_tile<T, BM, BN, BK, WM, WN, OP, VEC>(p, xcd_run(blockIdx.x, gridDim.x));
}

// Stream-K (FWD/WGRAD), with the schedule of lp_sk_kernel (sk_range /
// sk_next): sk_dp data-parallel tiles, then p.sk workgroups sharing the K
// tiles of the remaining tiles; the pieces of shared tiles are summed by
// sk_fixup_kernel in workgroup order.
template <typename T, int BM, int BN, int BK, int WM, int WN, int OP, bool VEC>
__global__ void __launch_bounds__(NT) igemm_sk_kernel(const cg_kparams p_arg) {
    const cg_kparams &p = kernarg_params(p_arg);
    SkRange r = sk_range(p);
    for (SkPiece q; sk_next(p, r, q, BM * BN);) {
        igemm_tile<T, BM, BN, BK, WM, WN, OP, VEC>(p, q.tile, q.k0, q.k1, q.piece);
        __syncthreads();  // LDS of this piece is free before the next one starts
    }
}

// Many independent problems of one configuration in one launch: item i
// belongs to the problem j with start[j] <= i < start[j+1] (nprob + 1
// entries).  Every start[j] is a multiple of 8 and a problem's range is its
// workgroup count rounded up to 8, so items 8k..8k+7 of a problem run on the
// 8 different XCDs and the XCD remap can be applied per problem: each XCD
// gets a contiguous run of every problem's tiles, and every problem is
// spread over all XCDs whatever the mix of problem costs.  Workgroup b runs
// items b, b + gridDim.x, ...; gridDim.x is the item count, or a multiple of
// 8 (persistent workgroups) so that all items of a workgroup keep its XCD.
template <typename T, int BM, int BN, int BK, int WM, int WN, int OP, bool VEC>
__global__ void __launch_bounds__(NT) igemm_grouped_kernel(const cg_kparams *__restrict__ plist,
                                                           const int *__restrict__ start, int nprob, int total) {
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const cg_kparams &p = plist[lo];
        const int lin = xcd_run(item - start[lo], start[lo + 1] - start[lo]);
        const int count = p.mtiles * p.ntiles * p.splits * p.ncls_h * p.ncls_w;
        __syncthreads();  // LDS reads of the previous item are complete
        if (lin < count) igemm_tile<T, BM, BN, BK, WM, WN, OP, VEC>(p, lin);
    }
}

// ---- instantiation helpers for the per-type launch files ----

// FP32 and BF16 serve all six operations; FP8 serves the forward pass.
template <typename T> constexpr bool op_enabled(int op) {
    return !std::is_same<T, uint8_t>::value || op == CG_FWD_NHWC || op == CG_FWD_NCHW;
}
// Does the VEC flag change any loader of this operation?
template <typename T> constexpr bool vec_relevant(int op) {
    return op == CG_FWD_NHWC || (op == CG_FWD_NCHW && fwd_b_kmajor<T>()) || op == CG_DGRAD_NHWC ||
           op == CG_DGRAD_NCHW || op == CG_WGRAD_NHWC;
}

// g == nullptr: one problem (kp), one workgroup per item.  g != nullptr: the
// grouped kernel over g's problems (kp only selects op and vec).  occ !=
// nullptr: report resident workgroups per CU instead of launching.
template <typename T, int BM, int BN, int BK, int WM, int WN, int OP, bool VEC>
inline int launch_igemm(const cg_kparams *kp, const cg_group *g, int *occ, hipStream_t st) {
    constexpr bool SKOK = OP == CG_FWD_NHWC || OP == CG_FWD_NCHW || OP == CG_WGRAD_NHWC || OP == CG_WGRAD_NCHW;
    if (occ) {
        if (kp->sk < 0) {  // occupancy of the stream-K kernel
            if constexpr (!SKOK) {
                *occ = 0;
                return 0;
            } else {
                return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    occ, reinterpret_cast<const void *>(&igemm_sk_kernel<T, BM, BN, BK, WM, WN, OP, VEC>), NT, 0);
            }
        }
        return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            occ, reinterpret_cast<const void *>(&igemm_grouped_kernel<T, BM, BN, BK, WM, WN, OP, VEC>), NT, 0);
    }
    if (kp->sk > 0 && !g) {
        if constexpr (!SKOK) {
            return (int)hipErrorInvalidValue;
        } else {
            hipLaunchKernelGGL((igemm_sk_kernel<T, BM, BN, BK, WM, WN, OP, VEC>),
                               dim3((unsigned)(kp->sk_dp + kp->sk)), dim3(NT), 0, st, *kp);
            return (int)hipGetLastError();
        }
    }
    if (g) {
        if (g->total <= 0) return 0;
        hipLaunchKernelGGL((igemm_grouped_kernel<T, BM, BN, BK, WM, WN, OP, VEC>), dim3((unsigned)g->grid), dim3(NT),
                           0, st, g->plist, g->start, g->nprob, g->total);
        return (int)hipGetLastError();
    }
    const long long tiles = (long long)kp->mtiles * kp->ntiles;
    const long long nb = tiles * kp->splits * kp->ncls_h * kp->ncls_w;
    hipLaunchKernelGGL((igemm_kernel<T, BM, BN, BK, WM, WN, OP, VEC>), dim3((unsigned)nb), dim3(NT), 0, st, *kp);
    return (int)hipGetLastError();
}

// NV: this configuration also has element-load (non-VEC) instantiations.
template <typename T, int BM, int BN, int BK, int WM, int WN, bool NV, int OP>
inline int launch_op(const cg_kparams *kp, const cg_group *g, int *occ, hipStream_t st) {
    if constexpr (!op_enabled<T>(OP)) {
        return (int)hipErrorInvalidValue;
    } else if constexpr (!vec_relevant<T>(OP)) {
        return launch_igemm<T, BM, BN, BK, WM, WN, OP, true>(kp, g, occ, st);
    } else {
        if (kp->vec) return launch_igemm<T, BM, BN, BK, WM, WN, OP, true>(kp, g, occ, st);
        if constexpr (NV) return launch_igemm<T, BM, BN, BK, WM, WN, OP, false>(kp, g, occ, st);
        return (int)hipErrorInvalidValue;
    }
}

}  // namespace cg
