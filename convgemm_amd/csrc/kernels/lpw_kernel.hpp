// lpw_kernel.hpp -- pipelined weight-gradient kernel (NHWC; FP32, BF16).
//
// dW = im2row(X)^T dY of src/convGemm.c:48-70 with trans 'T'
// (src/im2row_nhwc.c:pack_CB_nhwc_trans / pack_RB_nhwc_trans), as the GEMM
// C[(tap, c)][kn] = sum over output pixels p of X[p + tap][c] * dY[p][kn].
// Both operands are pixel-major in HBM: one pixel's channels (X) and its kn
// (dY) are contiguous.  So the operand tiles go from HBM straight into LDS
// as they lie -- a K step is KS pixels, an LDS row is one pixel's BM
// channels (A) or BN kn (B), each lane of a load instruction moves one
// 16-byte chunk (buffer_load ... lds), NS stages deep, with the explicit
// waits and fenced step regions of the forward pipelined kernel
// (lp_kernel.hpp) -- and the MFMA operands, which need K (pixels)
// contiguous per lane, are read from those row-major tiles transposed:
//   * BF16: ds_read_b64_tr_b16.  Within each 16-lane group, lane j receives
//     element j % 4 of the 8-byte row addressed by lane 4e + j / 4, as its
//     element e (e = 0..3).  Lane s of the group addresses pixel
//     k0 + s / 4 (+ 4 for the second read), channels m0 + 4 (s % 4) ..+3,
//     so lane j gets channel m0 + j at pixels k0 .. k0 + 3: two reads make
//     the 8 consecutive K of one 32x32x16 MFMA operand;
//   * FP32: the 32x32x2 MFMA takes one element per lane (row l % 32, K
//     l / 32): a plain ds_read_b32, consecutive lanes on consecutive
//     channels.
// Rows are not padded (the DMA writes 1 KB runs); the 16-byte chunk c of
// row r sits at position c ^ lpw_swz(r), which puts the reads of a half
// wave on distinct banks.  An M tile is one tap and BM channels (the
// general kernel's weight-gradient tiling), split-K divides the pixels.
// Padding taps, channels past C, kn past KN and pixels past P load from an
// out-of-range offset and read 0.  Each K step decomposes its pixels
// (b, oy, ox) once per load row.
#pragma once
#include "lp_kernel.hpp"

namespace cg {

typedef short i16x4 __attribute__((ext_vector_type(4)));

// LDS position of chunk c of row r: c ^ lpw_swz(r).  BF16 reads touch rows
// k..k+3 (k % 4 == 0) and four consecutive chunks per half wave; FP32 reads
// rows k and k+1 and eight consecutive chunks.  With 64 banks of 4 bytes
// these swizzles give each 16-byte slot of a half wave its own banks.
template <int ES, int CPR>
__device__ __forceinline__ int lpw_swz(int r) {
    if constexpr (ES == 4) return (r & 1) << 3;
    else if constexpr (CPR >= 16) return (r & 3) << 2;
    else return ((r >> 1) & 1) << 2;
}

template <typename T, int BM, int BN, int WM, int WN, int NS>
__device__ __attribute__((always_inline)) inline void lpw_tile(const cg_kparams &p, const int lin) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 2, "FP32 or BF16 operands");
    constexpr int ES = sizeof(T);
    constexpr int VEC = 16 / ES;               // elements per 16-byte chunk
    constexpr int KS = ES == 4 ? 32 : 64;      // pixels per K step
    constexpr int NW = WM * WN;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int TR = WTM / 32, TC = WTN / 32;
    constexpr int RA = BM * ES, RB = BN * ES;  // bytes of one pixel row of each operand
    constexpr int CPA = RA / 16, CPB = RB / 16;
    constexpr int PIA = 64 / CPA, PIB = 64 / CPB;  // pixel rows per load instruction
    constexpr int IA = KS / (PIA * NW), IB = KS / (PIB * NW);
    constexpr int SA = KS * RA;
    constexpr int STAGE = SA + KS * RB;
    constexpr int LPS = IA + IB;               // vm operations per wave per stage
    static_assert((NW == 4 || NW == 8) && TR >= 1 && TC >= 1 && WTM % 32 == 0 && WTN % 32 == 0, "wave tile");
    static_assert(CPA >= 8 && CPB >= 8 && CPA <= 64 && CPB <= 64, "rows of 128 to 1024 bytes");
    static_assert(IA >= 1 && IB >= 1 && KS % (PIA * NW) == 0 && KS % (PIB * NW) == 0, "load split");
    static_assert(NS >= 2 && NS <= 4 && (NS - 2) * LPS < 64, "stages");
    static_assert(ES == 2 || (CPA >= 16 && CPB >= 16), "FP32 swizzle needs rows of 16 chunks");

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
    const int tap = mt / p.kchunks;
    const int c0 = (mt - tap * p.kchunks) * BM;  // first channel of the tile
    const int kx = tap / p.KW, ky = tap - kx * p.KW;
    const int dh = kx * p.vd - p.vp, dw = ky * p.hd - p.hp;

    const __amdgpu_buffer_rsrc_t rX = make_rsrc(p.A, p.a_bytes);
    const __amdgpu_buffer_rsrc_t rY = make_rsrc(p.Bm, p.b_bytes);
    const int P = (int)p.P;
    const int HoWo = p.Ho * p.Wo;

    // ---- per-lane load rows: lane l of instruction j fills row (wave*I + j)*PI + l / CP,
    // position l % CP, with the global chunk position ^ swz(row) ----
    int a_row[IA], a_coff[IA];
    bool a_cok[IA];
#pragma unroll
    for (int j = 0; j < IA; j++) {
        const int r = (wave * IA + j) * PIA + lane / CPA;
        const int chunk = (lane % CPA) ^ lpw_swz<ES, CPA>(r);
        a_row[j] = r;
        a_cok[j] = c0 + chunk * VEC < p.C;
        a_coff[j] = (c0 + chunk * VEC) * ES;
    }
    int b_row[IB], b_noff[IB];
    bool b_nok[IB];
#pragma unroll
    for (int j = 0; j < IB; j++) {
        const int r = (wave * IB + j) * PIB + lane / CPB;
        const int chunk = (lane % CPB) ^ lpw_swz<ES, CPB>(r);
        const int n = nt * BN + chunk * VEC;
        b_row[j] = r;
        b_nok[j] = n < p.KN;
        b_noff[j] = n * ES;
    }

    // ---- reduction range (split-K over the K steps of KS pixels) ----
    const int KT = p.KT;
    const int per_split = (KT + p.splits - 1) / p.splits;
    const int kt_begin = min(KT, split * per_split);
    const int kt_end = min(KT, kt_begin + per_split);
    const int nsteps = kt_end - kt_begin;
    int i_kt = kt_begin;  // next K step to load

    auto stage_ptr = [&](auto s) __attribute__((always_inline)) -> char * {
        constexpr int S = decltype(s)::value;
        if constexpr (S == 0) return st0;
        else if constexpr (S == 1) return st1;
        else if constexpr (S == 2) return st2;
        else return st3;
    };

    // One K step: LPS vm operations per wave, X rows first; [J0, J1) of them.
    // The LDS destination is a plain 32-bit address (no alias scope: the
    // compiler adds no waits of its own); issues are never skipped.
    auto issue_part = [&](auto s, bool live, auto j0, auto j1) __attribute__((always_inline)) {
        constexpr int J0 = decltype(j0)::value, J1 = decltype(j1)::value;
        const unsigned base = (unsigned)(uintptr_t)(lds_void_t *)stage_ptr(s);
        const int p0 = i_kt * KS;
#pragma unroll
        for (int jj = J0; jj < J1; jj++) {
            if (jj < IA) {
                const int j = jj;
                const int pp = p0 + a_row[j];  // < 2^31: fdiv is exact past P as well
                const int b = fdiv(pp, x.howo);
                const int pix = pp - b * HoWo;
                const int oy = fdiv(pix, x.wo);
                const int ih = oy * p.vs + dh, iw = (pix - oy * p.Wo) * p.hs + dw;
                const bool ok = live && pp < P && a_cok[j] && (unsigned)ih < (unsigned)p.H &&
                                (unsigned)iw < (unsigned)p.W;
                const unsigned off =
                    ok ? (((unsigned)b * p.H + ih) * p.W + iw) * (unsigned)(p.C * ES) + a_coff[j] : OOB;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rX, (lds_void_t *)(uintptr_t)(base + (wave * IA + j) * 1024), 16, off, 0, 0, 0);
            } else {
                const int j = jj - IA;
                const int pp = p0 + b_row[j];
                const unsigned off =
                    (live && pp < P && b_nok[j]) ? (unsigned)pp * (unsigned)(p.KN * ES) + b_noff[j] : OOB;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rY, (lds_void_t *)(uintptr_t)(base + SA + (wave * IB + j) * 1024), 16, off, 0, 0, 0);
            }
        }
    };
    auto advance = [&]() __attribute__((always_inline)) { ++i_kt; };
    auto issue = [&](auto s, bool live) __attribute__((always_inline)) {
        issue_part(s, live, std::integral_constant<int, 0>{}, std::integral_constant<int, LPS>{});
        advance();
    };

    // ---- MFMA operand reads ----
    const int wm = wave % WM, wn = wave / WM;
    const int l32 = lane & 31, h = lane >> 5;
    // per-lane byte offsets (within a stage, before the group's K offset) of
    // this lane's operand element(s) in each 32-row block
    int aoff[TR], boff[TC];
    if constexpr (ES == 2) {
        const int G = lane >> 4, s = lane & 15;
        const int kl = 8 * (G >> 1) + (s >> 2);  // pixel within a group of 16 (first read)
        const int ml = 16 * (G & 1) + 4 * (s & 3);
        const int sa = lpw_swz<ES, CPA>(kl), sb = lpw_swz<ES, CPB>(kl);  // == those of kl + 4r + 16g
#pragma unroll
        for (int a = 0; a < TR; a++) {
            const int m = wm * WTM + a * 32 + ml;
            aoff[a] = kl * RA + (((m >> 3) ^ sa) << 4) + (m & 7) * 2;
        }
#pragma unroll
        for (int b = 0; b < TC; b++) {
            const int n = wn * WTN + b * 32 + ml;
            boff[b] = SA + kl * RB + (((n >> 3) ^ sb) << 4) + (n & 7) * 2;
        }
    } else {
        const int sa = lpw_swz<ES, CPA>(h), sb = lpw_swz<ES, CPB>(h);  // rows 2j + h
#pragma unroll
        for (int a = 0; a < TR; a++) {
            const int m = wm * WTM + a * 32 + l32;
            aoff[a] = h * RA + (((m >> 2) ^ sa) << 4) + (m & 3) * 4;
        }
#pragma unroll
        for (int b = 0; b < TC; b++) {
            const int n = wn * WTN + b * 32 + l32;
            boff[b] = SA + h * RB + (((n >> 2) ^ sb) << 4) + (n & 3) * 4;
        }
    }

    f32x16 acc[TR][TC];
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int b = 0; b < TC; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[a][b][v] = 0.f;

    // A stage holds NG = 4 groups: BF16 16 pixels (one MFMA of K = 16),
    // FP32 8 pixels (4 MFMAs of K = 2).  Two register sets alternate: the
    // reads of group g+1 are in flight while group g multiplies.
    constexpr int NG = 4;
    struct Ops16 {
        i32x4 a[TR], b[TC];
    };
    struct Ops32 {
        float a[TR][4], b[TC][4];
    };
    typedef typename std::conditional<ES == 2, Ops16, Ops32>::type Ops;
    auto ldg = [&](auto s, int g, Ops &r) __attribute__((always_inline)) {
        const char *base = stage_ptr(s);
        if constexpr (ES == 2) {
            typedef __attribute__((address_space(3))) i16x4 lds_v4;
            auto tr = [&](int off) __attribute__((always_inline)) -> i32x2 {
                const i16x4 v = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4 *)(base + off));
                return __builtin_bit_cast(i32x2, v);
            };
#pragma unroll
            for (int a = 0; a < TR; a++) {
                const i32x2 lo = tr(16 * g * RA + aoff[a]), hi = tr((16 * g + 4) * RA + aoff[a]);
                r.a[a] = i32x4{lo[0], lo[1], hi[0], hi[1]};
            }
#pragma unroll
            for (int b = 0; b < TC; b++) {
                const i32x2 lo = tr(16 * g * RB + boff[b]), hi = tr((16 * g + 4) * RB + boff[b]);
                r.b[b] = i32x4{lo[0], lo[1], hi[0], hi[1]};
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int kk = 2 * (4 * g + j);  // row of lane half 0
#pragma unroll
                for (int a = 0; a < TR; a++) r.a[a][j] = *(const float *)(base + kk * RA + aoff[a]);
#pragma unroll
                for (int b = 0; b < TC; b++) r.b[b][j] = *(const float *)(base + kk * RB + boff[b]);
            }
        }
    };
    auto mma = [&](const Ops &r) __attribute__((always_inline)) {
        if constexpr (ES == 2) {
#pragma unroll
            for (int a = 0; a < TR; a++)
#pragma unroll
                for (int b = 0; b < TC; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, r.a[a]),
                                                                        __builtin_bit_cast(bf16x8, r.b[b]), acc[a][b],
                                                                        0, 0, 0);
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int a = 0; a < TR; a++)
#pragma unroll
                    for (int b = 0; b < TC; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(r.a[a][j], r.b[b][j], acc[a][b], 0, 0, 0);
        }
    };

    Ops r0, r1;
    if constexpr (ES == 2) {
#pragma unroll
        for (int a = 0; a < TR; a++) r0.a[a] = i32x4{};
#pragma unroll
        for (int b = 0; b < TC; b++) r0.b[b] = i32x4{};
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) {
#pragma unroll
            for (int a = 0; a < TR; a++) r0.a[a][j] = 0.f;
#pragma unroll
            for (int b = 0; b < TC; b++) r0.b[b][j] = 0.f;
        }
    }
    r1 = r0;
    constexpr int PER = (LPS + NG - 1) / NG;  // loads per region
    // One stage in NG fenced regions, each reading one group's operands,
    // multiplying the previous group and sending a share of the loads of
    // step i + NS - 1 (lp_kernel.hpp's schedule); set 1 carries the last
    // group of a stage into the next.
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
    // Before step i: this wave's loads of step i are complete (the NS-2
    // later stages may still be in flight) and so are its LDS reads of step
    // i-1; the barrier then publishes stage i and frees stage (i-1) % NS.
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
        // tail: the last nsteps % NS steps sit in stages 0.. and are all
        // issued; nothing is overwritten any more
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

    // ---- write-back: rows = (tap, c), cols = kn ----
    const int mrow0 = tap * p.C + c0;  // logical row of the tile's first channel
    auto pos = [&](int a, int b, int v, int &m, int &n) __attribute__((always_inline)) {
        const int ml = wm * WTM + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
        n = nt * BN + wn * WTN + b * 32 + l32;
        m = mrow0 + ml;
        return c0 + ml < p.C && n < p.KN;
    };
    if (p.splits > 1) {  // raw split-K partials
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int b = 0; b < TC; b++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    int m, n;
                    if (pos(a, b, v, m, n)) p.ws[((long long)split * p.M + m) * p.N + n] = acc[a][b][v];
                }
        return;
    }
    if (p.beta == 0.0f) {  // nothing to load: one store per output
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int b = 0; b < TC; b++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    int m, n;
                    if (pos(a, b, v, m, n)) store_out<CG_WGRAD_NHWC>(p, x, m, n, acc[a][b][v]);
                }
        return;
    }
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int b = 0; b < TC; b++)
            store_block<CG_WGRAD_NHWC>(p, x, acc[a][b], nullptr, BN,
                                       [&](int v, int &m, int &n, int &nl) __attribute__((always_inline)) {
                                           nl = 0;
                                           return pos(a, b, v, m, n);
                                       });
}

template <typename T, int BM, int BN, int WM, int WN, int NS>
__global__ void __launch_bounds__(WM * WN * 64) lpw_kernel(const cg_kparams p) {
    lpw_tile<T, BM, BN, WM, WN, NS>(p, xcd_run(blockIdx.x, gridDim.x));
}

// Grouped form, with the item layout of igemm_grouped_kernel.
template <typename T, int BM, int BN, int WM, int WN, int NS>
__global__ void __launch_bounds__(WM * WN * 64) lpw_grouped_kernel(const cg_kparams *__restrict__ plist,
                                                                   const int *__restrict__ start, int nprob,
                                                                   int total) {
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
        if (lin < count) lpw_tile<T, BM, BN, WM, WN, NS>(p, lin);
    }
}

// Weight gradient NHWC with 16-byte groups of channels and kn; no stream-K form.
template <typename T, int BM, int BN, int WM, int WN, int NS>
inline int launch_lpw(const cg_kparams *kp, const cg_group *g, int *occ, hipStream_t st) {
    constexpr int NT = WM * WN * 64;
    if (kp->op != CG_WGRAD_NHWC || !kp->vec) return (int)hipErrorInvalidValue;
    if (occ) {
        if (kp->sk < 0) {  // the stream-K kernel's occupancy: there is none
            *occ = 0;
            return 0;
        }
        return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            occ, reinterpret_cast<const void *>(&lpw_grouped_kernel<T, BM, BN, WM, WN, NS>), NT, 0);
    }
    if (kp->sk > 0) return (int)hipErrorInvalidValue;
    if (g) {
        if (g->total <= 0) return 0;
        hipLaunchKernelGGL((lpw_grouped_kernel<T, BM, BN, WM, WN, NS>), dim3((unsigned)g->grid), dim3(NT), 0, st,
                           g->plist, g->start, g->nprob, g->total);
        return (int)hipGetLastError();
    }
    const long long nb = (long long)kp->mtiles * kp->ntiles * kp->splits;
    hipLaunchKernelGGL((lpw_kernel<T, BM, BN, WM, WN, NS>), dim3((unsigned)nb), dim3(NT), 0, st, *kp);
    return (int)hipGetLastError();
}

}  // namespace cg
