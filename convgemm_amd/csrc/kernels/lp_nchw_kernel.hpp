// lp_nchw_kernel.hpp -- pipelined implicit-GEMM forward kernel, NCHW, FP32.
//
// The forward NCHW convolution of src/convGemm.c:86-108 (im2col_nchw.c) on
// the pipeline of lp_kernel.hpp: operands go from HBM straight into NS LDS
// stages, every wait is explicit and each K step is split into fenced
// regions.  What changes is the activation side.  In NCHW a 16-byte run is
// 4 pixels of one channel, so the activation tile is kept K-major in LDS,
// [32 K rows][BM pixels], and filled with 4-byte loads: one instruction
// moves 64 consecutive pixels of one (channel, tap) row, for any stride,
// dilation and padding.  The per-lane offset is the pixel's (once per tap)
// and the channel's offset a scalar; a padded pixel or a row past C gets an
// offset out of range (the channel part stays below 2^31, so the sum does
// not wrap) and reads 0.  The MFMA multiplies W (rows = kn, 16-byte
// chunks, the repacked [kn][tap][c] order) by X (columns = pixels), so the
// lanes of the accumulator walk pixels and the NCHW store is coalesced.
#pragma once
#include "lp_kernel.hpp"

namespace cg {

template <int BM, int BN, int WM, int WN, int NS>
__device__ __attribute__((always_inline)) inline void lpn_tile(const cg_kparams &p, const int lin, const int sk_k0 = -1,
                                                         const int sk_k1 = 0, float *piece = nullptr) {
    constexpr int RB = 128;                    // W bytes per row per stage (32 fp32)
    constexpr int KS = 32;                     // K rows per stage
    constexpr int XP = BM * 4 + 32;            // X row pitch: lanes 32-63 read 4 rows down, 32 banks over
    constexpr int MH = BM / 64;                // 64-pixel halves of a row
    constexpr int WTN = BN / WN, WTM = BM / WM;  // wave tile: kn rows x pixel columns
    constexpr int TR = WTN / 32, TC = WTM / 32;
    constexpr int IX = KS * MH / 4;            // X loads per wave per stage
    constexpr int IW = BN / 32;                // W loads per wave per stage
    constexpr int LPS = IX + IW;
    constexpr int XB = KS * XP;                // X bytes per stage
    constexpr int STAGE = XB + BN * RB;
    static_assert(WM * WN == 4 && TR >= 1 && TC >= 1 && BM % 64 == 0 && BN % 32 == 0, "tile");
    static_assert(NS >= 2 && NS <= 4 && (NS - 2) * LPS < 64, "stages");

    __shared__ __attribute__((aligned(1024))) char st0[STAGE];
    __shared__ __attribute__((aligned(1024))) char st1[STAGE];
    __shared__ __attribute__((aligned(1024))) char st2[NS > 2 ? STAGE : 16];
    __shared__ __attribute__((aligned(1024))) char st3[NS > 3 ? STAGE : 16];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    const int tiles = p.mtiles * p.ntiles;
    const int tile = lin % tiles, split = lin / tiles;
    int mt, nt;
    tile_mn(p, tile, mt, nt);
    const Ctx x = make_ctx(p, 0);

    const __amdgpu_buffer_rsrc_t rX = make_rsrc(p.A, p.a_bytes);
    const __amdgpu_buffer_rsrc_t rW = make_rsrc(p.Bm, p.b_bytes);
    const unsigned HW4 = (unsigned)p.H * p.W * 4;  // bytes of one channel plane
    const int CB = p.C * 4;                        // bytes of one kn row of one tap

    // ---- per-lane pixel of each 64-pixel half ----
    int x_ih[MH], x_iw[MH];
    unsigned x_base[MH];
#pragma unroll
    for (int h2 = 0; h2 < MH; h2++) {
        const int m = mt * BM + h2 * 64 + lane;
        x_ih[h2] = -(1 << 29);
        x_iw[h2] = -(1 << 29);
        x_base[h2] = 0;
        if (m < (int)p.P) {
            const int b = fdiv(m, x.howo);
            const int pix = m - b * p.Ho * p.Wo;
            const int xo = fdiv(pix, x.wo);
            x_ih[h2] = xo * p.vs - p.vp;
            x_iw[h2] = (pix - xo * p.Wo) * p.hs - p.hp;
            x_base[h2] = (unsigned)b * (unsigned)p.C * HW4;
        }
    }
    // ---- W rows: lane l of instruction j fills row (wave*IW + j)*8 + l/8 ----
    const int lrow = lane >> 3, lpos = lane & 7;
    int w_off[IW], w_cb[IW];
    bool w_ok[IW];
#pragma unroll
    for (int j = 0; j < IW; j++) {
        const int r = (wave * IW + j) * 8 + lrow;
        const int chunk = lpos ^ ((r >> 1) & 7);
        w_cb[j] = chunk * 16;
        const int n = nt * BN + r;
        w_ok[j] = n < p.KN;
        w_off[j] = n * p.taps * CB + w_cb[j];
    }

    // ---- reduction range (split-K) ----
    const int KT = p.KT;
    const int per_split = (KT + p.splits - 1) / p.splits;
    // stream-K (sk_k0 >= 0): `lin` is the tile and K steps [sk_k0, sk_k1) are reduced
    const bool skm = sk_k0 >= 0;
    const int kt_begin = skm ? sk_k0 : min(KT, split * per_split);
    const int kt_end = skm ? sk_k1 : min(KT, kt_begin + per_split);
    const int nsteps = kt_end - kt_begin;

    int i_tap, i_cb;
    step_start(p, kt_begin, i_tap, i_cb);
    int i_kx = i_tap / p.KW;
    int i_ky = i_tap - i_kx * p.KW;

    auto stage_ptr = [&](auto s) __attribute__((always_inline)) -> char * {
        constexpr int S = decltype(s)::value;
        if constexpr (S == 0) return st0;
        else if constexpr (S == 1) return st1;
        else if constexpr (S == 2) return st2;
        else return st3;
    };

    // One K step: LPS vm operations per wave, X rows first.  The LDS
    // destination is a plain 32-bit address (no alias scope: the compiler
    // adds no waits of its own); issues are never skipped.
    auto issue_part = [&](auto s, bool live, auto j0, auto j1) __attribute__((always_inline)) {
        constexpr int J0 = decltype(j0)::value, J1 = decltype(j1)::value;
        const unsigned base = (unsigned)(uintptr_t)(lds_void_t *)stage_ptr(s);
        const int dh = i_kx * p.vd, dw = i_ky * p.hd;
        const int c0 = i_cb * KS;
        unsigned voff[MH];
#pragma unroll
        for (int h2 = 0; h2 < MH; h2++) {
            const int ih = x_ih[h2] + dh, iw = x_iw[h2] + dw;
            const bool ok = live && (unsigned)ih < (unsigned)p.H && (unsigned)iw < (unsigned)p.W;
            voff[h2] = ok ? x_base[h2] + (unsigned)(ih * p.W + iw) * 4u : OOB;
        }
        const int lim = CB - i_cb * RB;
        const int tapW = i_tap * CB + i_cb * RB;
#pragma unroll
        for (int jj = J0; jj < J1; jj++) {
            if (jj < IX) {
                const int q = wave * IX + jj;  // X instruction: row q / MH, half q % MH
                const int r = q / MH, h2 = q % MH;
                const int c = c0 + r;
                const bool cok = c < p.C;  // uniform: a row past C reads 0 in every lane
                const unsigned soff = cok ? (unsigned)c * HW4 : 0u;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rX, (lds_void_t *)(uintptr_t)(base + r * XP + h2 * 256),
                                                         4, cok ? voff[h2] : OOB, soff, 0, 0);
            } else {
                const int j = jj - IX;
                const bool ok = live && w_ok[j] && w_cb[j] < lim;
                const unsigned off = ok ? (unsigned)(w_off[j] + tapW) : OOB;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rW, (lds_void_t *)(uintptr_t)(base + XB + (wave * IW + j) * 1024), 16, off, 0, 0, 0);
            }
        }
    };
    auto advance = [&]() __attribute__((always_inline)) { step_next(p, i_tap, i_cb, i_kx, i_ky); };
    auto issue = [&](auto s, bool live) __attribute__((always_inline)) {
        issue_part(s, live, std::integral_constant<int, 0>{}, std::integral_constant<int, LPS>{});
        advance();
    };

    // ---- MFMA: A = W (rows kn), B = X (columns = pixels) ----
    const int wn = wave % WN, wm = wave / WN;
    const int l32 = lane & 31, h = lane >> 5;
    const int swz = (l32 >> 1) & 7;
    const int rowW = XB + (wn * WTN + l32) * RB;
    const int colX = (wm * WTM + l32) * 4;
    f32x16 acc[TR][TC];
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int b = 0; b < TC; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[a][b][v] = 0.f;

    // Group g of a stage covers K rows 8g..8g+7; lane half h supplies rows
    // 8g + 4h + j (j = 0..3) to the j-th MFMA of K = 2 on both operands: W
    // as one 16-byte chunk (chunk 2g + h of its row), X as four 4-byte reads.
    constexpr int NG = 4;
    struct Ops {
        i32x4 w[TR];
        float xv[TC][4];
    };
    auto ldg = [&](auto s, int g, Ops &r) __attribute__((always_inline)) {
        const char *base = stage_ptr(s);
        const int co = ((2 * g + h) ^ swz) * 16;
#pragma unroll
        for (int a = 0; a < TR; a++) r.w[a] = *(const i32x4 *)(base + rowW + a * 32 * RB + co);
        const int k0 = 8 * g + 4 * h;
#pragma unroll
        for (int b = 0; b < TC; b++)
#pragma unroll
            for (int j = 0; j < 4; j++) r.xv[b][j] = *(const float *)(base + (k0 + j) * XP + colX + b * 128);
    };
    auto mma = [&](const Ops &r) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int a = 0; a < TR; a++)
#pragma unroll
                for (int b = 0; b < TC; b++) {
                    // element j goes through a scalar: a bit_cast applied to a
                    // vector-element lvalue compiled to element 0 here
                    const int ew = r.w[a][j];
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(__builtin_bit_cast(float, ew), r.xv[b][j],
                                                                    acc[a][b], 0, 0, 0);
                }
    };

    Ops r0, r1;
#pragma unroll
    for (int a = 0; a < TR; a++) r0.w[a] = i32x4{};
#pragma unroll
    for (int b = 0; b < TC; b++)
#pragma unroll
        for (int j = 0; j < 4; j++) r0.xv[b][j] = 0.f;
    r1 = r0;
    constexpr int PER = (LPS + NG - 1) / NG;
    auto stage_mma = [&](auto s, auto issue_next, int i) __attribute__((always_inline)) {
        constexpr bool ISS = decltype(issue_next)::value;
        constexpr int SN = (decltype(s)::value + NS - 1) % NS;
        const bool live = i + NS - 1 < nsteps;
        auto region = [&](auto g) __attribute__((always_inline)) {
            constexpr int G = decltype(g)::value;
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (G % 2 == 0) {
                ldg(s, G, r0);
                mma(r1);
            } else {
                ldg(s, G, r1);
                mma(r0);
            }
            if constexpr (ISS && G * PER < LPS)
                issue_part(std::integral_constant<int, SN>{}, live, std::integral_constant<int, G * PER>{},
                           std::integral_constant<int, (G * PER + PER < LPS ? G * PER + PER : LPS)>{});
        };
        region(std::integral_constant<int, 0>{});
        region(std::integral_constant<int, 1>{});
        region(std::integral_constant<int, 2>{});
        region(std::integral_constant<int, 3>{});
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (ISS) advance();
    };
    auto sync = [&](auto pending) __attribute__((always_inline)) {
        wait_vm<decltype(pending)::value>();
        __builtin_amdgcn_s_waitcnt((0xF) | (0x7 << 4) | (0x0 << 8) | (0x3 << 14));  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
    };
    constexpr auto kInFlight = std::integral_constant<int, (NS - 2) * LPS>{};
    constexpr auto kYes = std::true_type{};
    constexpr auto kNo = std::false_type{};

    if (nsteps > 0) {
        issue(std::integral_constant<int, 0>{}, true);
        if constexpr (NS > 2) issue(std::integral_constant<int, 1>{}, 1 < nsteps);
        if constexpr (NS > 3) issue(std::integral_constant<int, (NS > 3 ? 2 : 0)>{}, 2 < nsteps);
        int i = 0;
        for (; i + NS <= nsteps; i += NS) {
            sync(kInFlight);
            stage_mma(std::integral_constant<int, 0>{}, kYes, i);
            sync(kInFlight);
            stage_mma(std::integral_constant<int, 1>{}, kYes, i + 1);
            if constexpr (NS > 2) {
                sync(kInFlight);
                stage_mma(std::integral_constant<int, (NS > 2 ? 2 : 0)>{}, kYes, i + 2);
            }
            if constexpr (NS > 3) {
                sync(kInFlight);
                stage_mma(std::integral_constant<int, (NS > 3 ? 3 : 0)>{}, kYes, i + 3);
            }
        }
        const int rem = nsteps - i;
        if (rem > 0) {
            sync(std::integral_constant<int, 0>{});
            stage_mma(std::integral_constant<int, 0>{}, kNo, 0);
            if (rem > 1) stage_mma(std::integral_constant<int, 1>{}, kNo, 0);
            if constexpr (NS > 3)
                if (rem > 2) stage_mma(std::integral_constant<int, (NS > 3 ? 2 : 0)>{}, kNo, 0);
        }
        mma(r1);
    }

    // ---- write-back: rows = kn, lanes = consecutive pixels ----
    if (piece || p.splits > 1) {  // raw partials: stream-K piece (tile-local) or split-K slice
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int b = 0; b < TC; b++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int nl = wn * WTN + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3), ml = wm * WTM + b * 32 + l32;
                    if (piece) {  // pixel-major, as the output: lanes store consecutive floats
                        piece[nl * BM + ml] = acc[a][b][v];
                        continue;
                    }
                    const int n = nt * BN + nl, m = mt * BM + ml;
                    if (m < x.Mrows && n < p.KN) p.ws[((long long)split * p.M + m) * p.N + n] = acc[a][b][v];
                }
        return;
    }
    if (!epi_staged<CG_FWD_NCHW>(p) && p.beta == 0.0f) {  // plain: nothing to load, one store per output
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int b = 0; b < TC; b++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int nl = wn * WTN + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3), ml = wm * WTM + b * 32 + l32;
                    const int m = mt * BM + ml, n = nt * BN + nl;
                    if (m < x.Mrows && n < p.KN) store_out<CG_FWD_NCHW>(p, x, m, n, acc[a][b][v]);
                }
        return;
    }
    // the fused epilogue: per-channel vectors staged in LDS (the pipeline stages are dead)
    float *s_epi = nullptr;
    if (epi_staged<CG_FWD_NCHW>(p)) {
        static_assert(5 * BN * sizeof(float) <= STAGE, "epilogue vectors fit one LDS stage");
        __syncthreads();
        s_epi = reinterpret_cast<float *>(st0);
        epi_stage(p, nt * BN, BN, s_epi, tid, WM * WN * 64);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int b = 0; b < TC; b++)
            store_block<CG_FWD_NCHW>(p, x, acc[a][b], s_epi, BN, [&](int v, int &m, int &n, int &nl) __attribute__((always_inline)) {
                nl = wn * WTN + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
                const int ml = wm * WTM + b * 32 + l32;
                n = nt * BN + nl;
                m = mt * BM + ml;
                return m < x.Mrows && n < p.KN;
            });
    if (s_epi) __syncthreads();  // the LDS is rewritten by the next item of a persistent launch
}

template <int BM, int BN, int WM, int WN, int NS>
__global__ void __launch_bounds__(LP_NT) lpn_kernel(const cg_kparams p) {
    lpn_tile<BM, BN, WM, WN, NS>(p, xcd_run(blockIdx.x, gridDim.x));
}

// Stream-K, with the schedule of lp_sk_kernel: sk_dp data-parallel tiles,
// then p.sk workgroups sharing the K steps of the remaining tiles.
template <int BM, int BN, int WM, int WN, int NS>
__global__ void __launch_bounds__(LP_NT) lpn_sk_kernel(const cg_kparams p_arg) {
    const cg_kparams &p = kernarg_params(p_arg);
    SkRange r = sk_range(p);
    for (SkPiece q; sk_next(p, r, q, BM * BN);) {
        lpn_tile<BM, BN, WM, WN, NS>(p, q.tile, q.k0, q.k1, q.piece);
        __syncthreads();  // loads, LDS reads and stores of this piece are complete
    }
}

template <int BM, int BN, int WM, int WN, int NS>
__global__ void __launch_bounds__(LP_NT) lpn_grouped_kernel(const cg_kparams *__restrict__ plist,
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
        const int count = p.mtiles * p.ntiles * p.splits;
        __syncthreads();  // every LDS read and load of the previous item is complete
        if (lin < count) lpn_tile<BM, BN, WM, WN, NS>(p, lin);
    }
}

// Forward NCHW, FP32, weights in 16-byte chunks (C % 4 == 0, aligned).
template <int BM, int BN, int WM, int WN, int NS>
inline int launch_lpn(const cg_kparams *kp, const cg_group *g, int *occ, hipStream_t st) {
    if (kp->op != CG_FWD_NCHW || !kp->vec) return (int)hipErrorInvalidValue;
    if (occ) {
        const void *k = kp->sk < 0 ? reinterpret_cast<const void *>(&lpn_sk_kernel<BM, BN, WM, WN, NS>)
                                   : reinterpret_cast<const void *>(&lpn_grouped_kernel<BM, BN, WM, WN, NS>);
        return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k, LP_NT, 0);
    }
    if (kp->sk > 0 && !g) {
        hipLaunchKernelGGL((lpn_sk_kernel<BM, BN, WM, WN, NS>), dim3((unsigned)(kp->sk_dp + kp->sk)), dim3(LP_NT), 0,
                           st, *kp);
        return (int)hipGetLastError();
    }
    if (g) {
        if (g->total <= 0) return 0;
        hipLaunchKernelGGL((lpn_grouped_kernel<BM, BN, WM, WN, NS>), dim3((unsigned)g->grid), dim3(LP_NT), 0, st,
                           g->plist, g->start, g->nprob, g->total);
        return (int)hipGetLastError();
    }
    const long long nb = (long long)kp->mtiles * kp->ntiles * kp->splits;
    hipLaunchKernelGGL((lpn_kernel<BM, BN, WM, WN, NS>), dim3((unsigned)nb), dim3(LP_NT), 0, st, *kp);
    return (int)hipGetLastError();
}

}  // namespace cg
