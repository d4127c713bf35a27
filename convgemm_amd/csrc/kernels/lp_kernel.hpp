// lp_kernel.hpp -- pipelined implicit-GEMM forward kernel (NHWC, BF16/FP8).
//
// The forward NHWC convolution of src/convGemm.c:48-70 for 2-byte and
// 1-byte operands, where one K step of the general kernel (8 MFMAs per wave)
// is too short to cover the latency of its gathered loads.  Here:
//   * every K step moves 128 bytes of each operand row (64 bf16 / 128 fp8
//     channels of one tap): 16 MFMAs per wave for a 64x64 wave tile;
//   * operands go from HBM straight into LDS (buffer_load_dwordx4 ... lds:
//     each wave instruction fills 8 rows x 128 B), no register staging, so
//     the pipeline is NS LDS stages deep (NS-1 steps of loads in flight);
//   * stage buffers are separate LDS objects and every wait is explicit
//     (vmcnt for the stage about to be read, then s_barrier), so the
//     compiler's memory model does not drain the pipeline at each barrier;
//   * rows are 128 B with the 16-byte chunk c of row r stored at position
//     c ^ ((r >> 1) & 7): the MFMA operand reads (16 lanes = 16 consecutive
//     rows, one chunk) hit all 64 banks once.
// Weights are read in the repacked [kn][tap][c] order.  Padding taps,
// channels past C and rows past M/N load from an out-of-range offset and
// read 0.  The write-back (split-K partials or the fused epilogue) is the
// general kernel's.
#pragma once
#include <type_traits>

#include "igemm_common.hpp"

namespace cg {

constexpr int LP_NT = 256;

// s_waitcnt immediate: wait until at most n vector-memory operations of this
// wave are outstanding (expcnt and lgkmcnt left at their no-wait maxima).
template <int N> __device__ __attribute__((always_inline)) inline void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

typedef __attribute__((address_space(3))) void lds_void_t;

template <typename T, int BM, int BN, int WM, int WN, int NS>
__device__ __attribute__((always_inline)) inline void lp_tile(const cg_kparams &p, const int lin) {
    static_assert(sizeof(T) == 2 || sizeof(T) == 1, "BF16 or FP8 operands");
    constexpr int ES = sizeof(T);
    constexpr int RB = 128;                 // bytes per row per stage
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int TR = WTM / 32, TC = WTN / 32;
    constexpr int IA = BM / 32, IB = BN / 32;  // load instructions per wave per stage
    constexpr int STAGE = (BM + BN) * RB;
    constexpr int LPS = IA + IB;               // vm operations per wave per stage
    static_assert(WM * WN == 4 && TR >= 1 && TC >= 1 && WTM % 32 == 0 && WTN % 32 == 0, "wave tile");
    static_assert(BM % 32 == 0 && BN % 32 == 0 && NS >= 2 && NS <= 4, "tile");
    static_assert((NS - 2) * LPS < 64, "vmcnt");

    __shared__ __attribute__((aligned(1024))) char st0[STAGE];
    __shared__ __attribute__((aligned(1024))) char st1[STAGE];
    __shared__ __attribute__((aligned(1024))) char st2[NS > 2 ? STAGE : 16];
    __shared__ __attribute__((aligned(1024))) char st3[NS > 3 ? STAGE : 16];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: LDS addresses go to M0

    const int tiles = p.mtiles * p.ntiles;
    const int tile = lin % tiles, split = lin / tiles;
    const int nt = tile % p.ntiles, mt = tile / p.ntiles;
    const Ctx x = make_ctx(p, 0);

    const __amdgpu_buffer_rsrc_t rA = make_rsrc(p.A, p.a_bytes);
    const __amdgpu_buffer_rsrc_t rB = make_rsrc(p.Bm, p.b_bytes);
    const int CB = p.C * ES;  // bytes of one pixel's channels (a multiple of 16)

    // ---- per-lane load rows: lane l of instruction j fills LDS row
    // (wave*I + j)*8 + l/8, position l%8, with global chunk pos ^ swz(row) ----
    const int lrow = lane >> 3, lpos = lane & 7;
    int a_ih[IA], a_iw[IA], a_off[IA], a_cb[IA];
#pragma unroll
    for (int j = 0; j < IA; j++) {
        const int r = (wave * IA + j) * 8 + lrow;
        const int chunk = lpos ^ ((r >> 1) & 7);
        a_cb[j] = chunk * 16;
        const int m = mt * BM + r;
        a_ih[j] = -(1 << 29);
        a_iw[j] = -(1 << 29);
        a_off[j] = 0;
        if (m < (int)p.P) {
            const int b = fdiv(m, x.howo);
            const int pix = m - b * p.Ho * p.Wo;
            const int xo = fdiv(pix, x.wo);
            const int ih0 = xo * p.vs - p.vp, iw0 = (pix - xo * p.Wo) * p.hs - p.hp;
            a_ih[j] = ih0;
            a_iw[j] = iw0;
            a_off[j] = ((b * p.H + ih0) * p.W + iw0) * CB + a_cb[j];
        }
    }
    int b_off[IB], b_cb[IB];
    bool b_ok[IB];
#pragma unroll
    for (int j = 0; j < IB; j++) {
        const int r = (wave * IB + j) * 8 + lrow;
        const int chunk = lpos ^ ((r >> 1) & 7);
        b_cb[j] = chunk * 16;
        const int n = nt * BN + r;
        b_ok[j] = n < p.KN;
        b_off[j] = n * p.taps * CB + b_cb[j];
    }

    // ---- reduction range (split-K) ----
    const int KT = p.KT;
    const int per_split = (KT + p.splits - 1) / p.splits;
    const int kt_begin = min(KT, split * per_split);
    const int kt_end = min(KT, kt_begin + per_split);
    const int nsteps = kt_end - kt_begin;

    // issue iterator: the next K step to load (tap, channel block)
    int i_tap = kt_begin / p.kchunks;
    int i_cb = kt_begin - i_tap * p.kchunks;
    int i_kx = i_tap / p.KW;
    int i_ky = i_tap - i_kx * p.KW;

    auto stage_ptr = [&](auto s) __attribute__((always_inline)) -> char * {
        constexpr int S = decltype(s)::value;
        if constexpr (S == 0) return st0;
        else if constexpr (S == 1) return st1;
        else if constexpr (S == 2) return st2;
        else return st3;
    };

    // gather one K step into a stage (LPS vm operations per wave); live ==
    // false loads zeros (out-of-range offsets).  Issues are never skipped: a
    // conditional issue makes the compiler drain every load before the next
    // LDS read.
    auto issue = [&](auto s, bool live) __attribute__((always_inline)) {
        char *base = stage_ptr(s);
        const int dh = i_kx * p.vd, dw = i_ky * p.hd;
        const int tapA = (dh * p.W + dw) * CB + i_cb * RB;
        const int lim = CB - i_cb * RB;  // channel bytes left in this block
        const int tapB = i_tap * CB + i_cb * RB;
#pragma unroll
        for (int j = 0; j < IA; j++) {
            const int ih = a_ih[j] + dh, iw = a_iw[j] + dw;
            const bool ok = live && (unsigned)ih < (unsigned)p.H && (unsigned)iw < (unsigned)p.W && a_cb[j] < lim;
            const unsigned off = ok ? (unsigned)(a_off[j] + tapA) : OOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rA, (lds_void_t *)(base + (wave * IA + j) * 1024), 16, off, 0,
                                                     0, 0);
        }
#pragma unroll
        for (int j = 0; j < IB; j++) {
            const bool ok = live && b_ok[j] && b_cb[j] < lim;
            const unsigned off = ok ? (unsigned)(b_off[j] + tapB) : OOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rB, (lds_void_t *)(base + BM * RB + (wave * IB + j) * 1024),
                                                     16, off, 0, 0, 0);
        }
        if (++i_cb == p.kchunks) {
            i_cb = 0;
            ++i_tap;
            if (++i_ky == p.KW) { i_ky = 0; ++i_kx; }
        }
    };

    // ---- MFMA ----
    const int wm = wave % WM, wn = wave / WM;
    const int l32 = lane & 31, h = lane >> 5;
    const int swz = (l32 >> 1) & 7;  // rows handled by this lane are l32 mod 32
    const int rowA = (wm * WTM + l32) * RB, rowB = BM * RB + (wn * WTN + l32) * RB;
    f32x16 acc[TR][TC];
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int b = 0; b < TC; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[a][b][v] = 0.f;

    // MFMA groups [G0, G1) of a stage (4 groups of 16 bf16 / 2 of 64 fp8)
    constexpr int NG = ES == 2 ? 4 : 2;
    auto compute = [&](auto s, auto g0, auto g1) __attribute__((always_inline)) {
        constexpr int G0 = decltype(g0)::value, G1 = decltype(g1)::value;
        const char *base = stage_ptr(s);
        if constexpr (ES == 2) {
#pragma unroll
            for (int g = G0; g < G1; g++) {  // 16 bf16 per group
                const int c = 2 * g + h;
                const int co = (c ^ swz) * 16;
                u16x8 ar[TR], bc[TC];
#pragma unroll
                for (int a = 0; a < TR; a++) ar[a] = *(const u16x8 *)(base + rowA + a * 32 * RB + co);
#pragma unroll
                for (int b = 0; b < TC; b++) bc[b] = *(const u16x8 *)(base + rowB + b * 32 * RB + co);
#pragma unroll
                for (int a = 0; a < TR; a++)
#pragma unroll
                    for (int b = 0; b < TC; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, ar[a]), __builtin_bit_cast(bf16x8, bc[b]), acc[a][b], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int g = G0; g < G1; g++) {  // 64 fp8 per group
                const int c0 = 4 * g + 2 * h;
                const int co0 = (c0 ^ swz) * 16, co1 = ((c0 + 1) ^ swz) * 16;
                i32x8 ar[TR], bc[TC];
#pragma unroll
                for (int a = 0; a < TR; a++) {
                    const i32x4 lo = *(const i32x4 *)(base + rowA + a * 32 * RB + co0);
                    const i32x4 hi = *(const i32x4 *)(base + rowA + a * 32 * RB + co1);
                    ar[a] = i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                }
#pragma unroll
                for (int b = 0; b < TC; b++) {
                    const i32x4 lo = *(const i32x4 *)(base + rowB + b * 32 * RB + co0);
                    const i32x4 hi = *(const i32x4 *)(base + rowB + b * 32 * RB + co1);
                    bc[b] = i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                }
#pragma unroll
                for (int a = 0; a < TR; a++)
#pragma unroll
                    for (int b = 0; b < TC; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(ar[a], bc[b], acc[a][b], 0, 0, 0,
                                                                                    127, 0, 127);
            }
        }
    };

    // Step i reads stage i % NS.  Before it: this wave's loads of step i are
    // complete (the NS-2 later stages may still be in flight), its LDS reads
    // of step i-1 are complete, then the barrier makes every wave's part
    // visible and frees stage (i-1) % NS for the loads of step i + NS - 1.
    auto step = [&](auto s, int i) __attribute__((always_inline)) {
        constexpr int S = decltype(s)::value;
        wait_vm<(NS - 2) * LPS>();
        __builtin_amdgcn_s_waitcnt((0xF) | (0x7 << 4) | (0x0 << 8) | (0x3 << 14));  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        // the first MFMA group goes ahead of the loads' address arithmetic
        compute(s, std::integral_constant<int, 0>{}, std::integral_constant<int, 1>{});
        issue(std::integral_constant<int, (S + NS - 1) % NS>{}, i + NS - 1 < nsteps);
        compute(s, std::integral_constant<int, 1>{}, std::integral_constant<int, NG>{});
    };

    if (nsteps > 0) {
        issue(std::integral_constant<int, 0>{}, true);
        if constexpr (NS > 2) issue(std::integral_constant<int, 1>{}, 1 < nsteps);
        if constexpr (NS > 3) issue(std::integral_constant<int, (NS > 3 ? 2 : 0)>{}, 2 < nsteps);
        int i = 0;
        for (; i + NS <= nsteps; i += NS) {
            step(std::integral_constant<int, 0>{}, i);
            step(std::integral_constant<int, 1>{}, i + 1);
            if constexpr (NS > 2) step(std::integral_constant<int, (NS > 2 ? 2 : 0)>{}, i + 2);
            if constexpr (NS > 3) step(std::integral_constant<int, (NS > 3 ? 3 : 0)>{}, i + 3);
        }
        // tail: the last nsteps % NS steps sit in stages 0.. and are all
        // issued; nothing is overwritten any more
        const int rem = nsteps - i;
        if (rem > 0) {
            wait_vm<0>();
            __builtin_amdgcn_s_waitcnt((0xF) | (0x7 << 4) | (0x0 << 8) | (0x3 << 14));
            __builtin_amdgcn_s_barrier();
            constexpr auto g0 = std::integral_constant<int, 0>{};
            constexpr auto gN = std::integral_constant<int, NG>{};
            compute(std::integral_constant<int, 0>{}, g0, gN);
            if (rem > 1) compute(std::integral_constant<int, 1>{}, g0, gN);
            if constexpr (NS > 3)
                if (rem > 2) compute(std::integral_constant<int, (NS > 3 ? 2 : 0)>{}, g0, gN);
        }
    }

    // ---- write-back (FWD NHWC: rows = output pixels, cols = kn) ----
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int b = 0; b < TC; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int m = mt * BM + wm * WTM + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
                const int n = nt * BN + wn * WTN + b * 32 + l32;
                if (m >= x.Mrows || n >= p.KN) continue;
                if (p.splits > 1)
                    p.ws[((long long)split * p.M + m) * p.N + n] = acc[a][b][v];
                else
                    store_out<CG_FWD_NHWC>(p, x, m, n, acc[a][b][v]);
            }
}

__device__ inline int lp_xcd_remap(int bid, int nb) {
    const int xcd = bid & 7, per = nb >> 3, rem = nb & 7;
    return xcd * per + min(xcd, rem) + (bid >> 3);
}

template <typename T, int BM, int BN, int WM, int WN, int NS>
__global__ void __launch_bounds__(LP_NT) lp_kernel(const cg_kparams p) {
    lp_tile<T, BM, BN, WM, WN, NS>(p, lp_xcd_remap(blockIdx.x, gridDim.x));
}

// Grouped form, with the item layout of igemm_grouped_kernel.
template <typename T, int BM, int BN, int WM, int WN, int NS>
__global__ void __launch_bounds__(LP_NT) lp_grouped_kernel(const cg_kparams *__restrict__ plist,
                                                           const int *__restrict__ start, int nprob, int total) {
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const cg_kparams &p = plist[lo];
        const int lin = lp_xcd_remap(item - start[lo], start[lo + 1] - start[lo]);
        const int count = p.mtiles * p.ntiles * p.splits;
        __syncthreads();  // every LDS read and load of the previous item is complete
        if (lin < count) lp_tile<T, BM, BN, WM, WN, NS>(p, lin);
    }
}

// Forward NHWC with 16-byte channel chunks only.
template <typename T, int BM, int BN, int WM, int WN, int NS>
inline int launch_lp(const cg_kparams *kp, const cg_group *g, int *occ, hipStream_t st) {
    if (kp->op != CG_FWD_NHWC || !kp->vec) return (int)hipErrorInvalidValue;
    if (occ) {
        return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            occ, reinterpret_cast<const void *>(&lp_grouped_kernel<T, BM, BN, WM, WN, NS>), LP_NT, 0);
    }
    if (g) {
        if (g->total <= 0) return 0;
        hipLaunchKernelGGL((lp_grouped_kernel<T, BM, BN, WM, WN, NS>), dim3((unsigned)g->grid), dim3(LP_NT), 0, st,
                           g->plist, g->start, g->nprob, g->total);
        return (int)hipGetLastError();
    }
    const long long nb = (long long)kp->mtiles * kp->ntiles * kp->splits;
    hipLaunchKernelGGL((lp_kernel<T, BM, BN, WM, WN, NS>), dim3((unsigned)nb), dim3(LP_NT), 0, st, *kp);
    return (int)hipGetLastError();
}

}  // namespace cg
