// lp_kernel.hpp -- pipelined implicit-GEMM forward kernel (NHWC; FP32, BF16, FP8).
//
// The forward NHWC convolution of src/convGemm.c:48-70 with 16-byte operand
// chunks, where one K step of the general kernel is too short to cover the
// latency of its gathered loads, or its register staging costs too many
// instructions.  Here:
//   * every K step moves 128 bytes of each operand row (32 fp32 / 64 bf16 /
//     128 fp8 channels of one tap): for a 64x64 wave tile, 64 / 16 / 8 MFMAs
//     per wave;
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

constexpr int LP_NT = 256;  // threads of a 4-wave workgroup (the NCHW kernel); the NHWC kernel runs 4 or 8 waves

// s_waitcnt immediate: wait until at most n vector-memory operations of this
// wave are outstanding (expcnt and lgkmcnt left at their no-wait maxima).
template <int N> __device__ __attribute__((always_inline)) inline void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

typedef __attribute__((address_space(3))) void lds_void_t;

// Order of the K steps of a forward tile, a step being (tap, block of
// channels): the KH x KW taps of one channel block run back to back, so the
// input rows one tap gathers are gathered again by the next taps while they
// are still in L2 (with the taps of one block outermost instead, the NCHW
// 3x3 stride-2 layer at batch 32 fetched 4.4x its compulsory bytes).  A
// split-K slice or a stream-K range is a contiguous run of steps.
__device__ __attribute__((always_inline)) inline void step_start(const cg_kparams &p, int s, int &tap, int &cb) {
    cb = s / p.taps;
    tap = s - cb * p.taps;
}
// Advance (tap, cb) and the tap's (kx, ky) = (tap / KW, tap mod KW) by one step.
__device__ __attribute__((always_inline)) inline void step_next(const cg_kparams &p, int &tap, int &cb, int &kx,
                                                                int &ky) {
    ++tap;
    if (++ky == p.KW) {
        ky = 0;
        if (++kx == p.KH) {
            kx = 0;
            tap = 0;
            ++cb;
        }
    }
}

// EXT: the NS stages live at ext + s * STAGE (a kernel that runs several tile
// shapes gives them one LDS buffer of the largest size); otherwise in LDS
// objects of the tile's own.
// RAWB: p.Bm holds the caller's weights, [c][kh][kw][kn], read in place.  A
// stage's weight tile is then KSW rows, one per channel of the step, each
// holding the tile's BN kn (chunk c of row r at c ^ rw_swz(r)); the MFMA
// reads it transposed: BF16 with ds_read_b64_tr_b16 (lane j of a 16-lane
// group gets column m0 + j at 4 consecutive rows, as in the weight-gradient
// kernel), FP8 with ds_read_b64_tr_b8 (lane j of a 16-lane group gets
// column m0 + j at 8 consecutive rows: lanes 2e and 2e + 1 address row e,
// columns m0..m0+7 and m0+8..m0+15), FP32 one 4-byte element per MFMA.  The
// K order is the activations': lane half h, group g, element e is channel
// 16g + 8h + e (BF16), 64g + 32h + e (FP8) or 8g + 4h + e (FP32) of the step.
// The FP8 swizzles put the 8 rows x 16 bytes of one 16-lane read on 32
// distinct banks, and the two chunks of a half-wave on 64.
template <int ES, int CPW>
__device__ __forceinline__ int rw_swz(int r) {
    if constexpr (ES == 4) return ((r >> 2) & 1) << 3;  // rows 8g + j and 8g + 4 + j on distinct bank halves
    else if constexpr (ES == 1) return CPW >= 8 ? (r & 7) : ((r >> 1) & 3);
    else if constexpr (CPW >= 16) return (r & 3) << 2;
    else return ((r >> 1) & 1) << 2;
}

// In-place weights, load side: instruction j of a wave fills the weight rows
// (wave * IB + j) * PIW + lane / CPW of a stage (PIW rows of CPW 16-byte
// chunks per instruction).  b_cb[j] is the row (the channel within the K
// step), b_off[j] the byte offset of its chunk without the K step's part.
template <int ES, int BN, int IB>
__device__ __forceinline__ void raw_rows(const cg_kparams &p, int wave, int lane, int nt, int (&b_off)[IB],
                                         int (&b_cb)[IB], bool (&b_ok)[IB]) {
    constexpr int CPW = BN * ES / 16, PIW = 64 / CPW;
#pragma unroll
    for (int j = 0; j < IB; j++) {
        const int r = (wave * IB + j) * PIW + lane / CPW;
        const int chunk = (lane % CPW) ^ rw_swz<ES, CPW>(r);
        const int n = nt * BN + chunk * (16 / ES);
        b_cb[j] = r;
        b_ok[j] = n < p.KN;
        b_off[j] = (int)(((unsigned)r * (unsigned)p.taps * (unsigned)p.KN + (unsigned)n) * ES);
    }
}

// Is row b_cb of K step (channel block cb) a channel (rows past C read 0)?
template <int ES>
__device__ __forceinline__ bool raw_live(const cg_kparams &p, int b_cb, int cb) {
    return b_cb < p.C - cb * (128 / ES);
}

// Global byte offset of the in-place weight load with base b_off in K step (cb, tap).
template <int ES>
__device__ __forceinline__ unsigned raw_off(const cg_kparams &p, int b_off, int cb, int tap) {
    return (unsigned)b_off + ((unsigned)(cb * (128 / ES)) * (unsigned)p.taps + (unsigned)tap) * (unsigned)p.KN * ES;
}

// In-place weights, read side: this lane's byte offset, within a weight
// stage, for each 32-column block b of the wave's columns col0 + 32b ..
template <int ES, int BN, int TC>
__device__ __forceinline__ void raw_woff(int lane, int col0, int (&woff)[TC]) {
    constexpr int RWB = BN * ES, CPW = RWB / 16;
    if constexpr (ES == 2) {
        const int G = lane >> 4, s16 = lane & 15;
        const int kl = 8 * (G >> 1) + (s16 >> 2);  // first row of the group's first read
        const int ml = 16 * (G & 1) + 4 * (s16 & 3);
        const int sw = rw_swz<ES, CPW>(kl);       // == that of kl + 4 and kl + 16g
#pragma unroll
        for (int b = 0; b < TC; b++) {
            const int n = col0 + b * 32 + ml;
            woff[b] = kl * RWB + (((n >> 3) ^ sw) << 4) + (n & 7) * 2;
        }
    } else if constexpr (ES == 1) {
        const int G = lane >> 4, s16 = lane & 15;
        const int kl = 32 * (G >> 1) + (s16 >> 1);  // row of read 0 of group 0
        const int sw = rw_swz<ES, CPW>(kl);         // == that of kl + 64g + 8q
#pragma unroll
        for (int b = 0; b < TC; b++) {
            const int c = ((col0 + b * 32) >> 4) + (G & 1);
            woff[b] = kl * RWB + ((c ^ sw) << 4) + 8 * (s16 & 1);
        }
    } else {
        const int l32 = lane & 31, h = lane >> 5;
#pragma unroll
        for (int b = 0; b < TC; b++) {
            const int n = col0 + b * 32 + l32;
            woff[b] = 4 * h * RWB + (((n >> 2) ^ (h << 3)) << 4) + (n & 3) * 4;
        }
    }
}

// The MFMA B operands of read group g from a weight stage at wst (the K
// order of the activations' operand).
template <int ES, int BN, int TC, typename OpV>
__device__ __forceinline__ void raw_read(const char *wst, const int (&woff)[TC], int g, OpV (&b)[TC]) {
    constexpr int RWB = BN * ES;
    if constexpr (ES == 2) {
        auto tr = [&](int off) __attribute__((always_inline)) -> i32x2 {
            typedef short s16x4 __attribute__((ext_vector_type(4)));
            typedef __attribute__((address_space(3))) s16x4 lds_s16x4;
            return __builtin_bit_cast(i32x2, __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)(wst + off)));
        };
#pragma unroll
        for (int c = 0; c < TC; c++) {
            const i32x2 lo = tr(woff[c] + 16 * g * RWB), hi = tr(woff[c] + (16 * g + 4) * RWB);
            b[c] = i32x4{lo[0], lo[1], hi[0], hi[1]};
        }
    } else if constexpr (ES == 1) {
        auto tr = [&](int off) __attribute__((always_inline)) -> i32x2 {
            typedef __attribute__((address_space(3))) i32x2 lds_i32x2;
            return __builtin_amdgcn_ds_read_tr8_b64_v2i32((lds_i32x2 *)(wst + off));
        };
#pragma unroll
        for (int c = 0; c < TC; c++) {
            const int o = woff[c] + 64 * g * RWB;
            const i32x2 q0 = tr(o), q1 = tr(o + 8 * RWB), q2 = tr(o + 16 * RWB), q3 = tr(o + 24 * RWB);
            b[c] = i32x8{q0[0], q0[1], q1[0], q1[1], q2[0], q2[1], q3[0], q3[1]};
        }
    } else {
#pragma unroll
        for (int c = 0; c < TC; c++)
#pragma unroll
            for (int j = 0; j < 4; j++) b[c][j] = *(const int *)(wst + woff[c] + (8 * g + j) * RWB);
    }
}

template <typename T, int BM, int BN, int WM, int WN, int NS, bool EXT = false, bool RAWB = false>
__device__ __attribute__((always_inline)) inline void lp_tile(const cg_kparams &p, const int lin, const int sk_k0 = -1,
                                                         const int sk_k1 = 0, float *piece = nullptr,
                                                         char *ext = nullptr) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 2 || sizeof(T) == 1, "FP32, BF16 or FP8 operands");
    constexpr int ES = sizeof(T);
    constexpr int RB = 128;                 // bytes per row per stage
    constexpr int NW = WM * WN;                // 4 or 8 waves
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int TR = WTM / 32, TC = WTN / 32;
    constexpr int IA = BM / (8 * NW), IB = BN / (8 * NW);  // load instructions per wave per stage
    constexpr int STAGE = (BM + BN) * RB;
    constexpr int LPS = IA + IB;               // vm operations per wave per stage
    static_assert((NW == 4 || NW == 8) && TR >= 1 && TC >= 1 && WTM % 32 == 0 && WTN % 32 == 0, "wave tile");
    static_assert(BM % (8 * NW) == 0 && BN % (8 * NW) == 0 && NS >= 2 && NS <= 4, "tile");
    static_assert((NS - 2) * LPS < 64, "vmcnt");

    __shared__ __attribute__((aligned(1024))) char st0[STAGE];
    __shared__ __attribute__((aligned(1024))) char st1[STAGE];
    __shared__ __attribute__((aligned(1024))) char st2[NS > 2 ? STAGE : 16];
    __shared__ __attribute__((aligned(1024))) char st3[NS > 3 ? STAGE : 16];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: LDS addresses go to M0

    const int tiles = p.mtiles * p.ntiles;
    const int tile = lin % tiles, split = lin / tiles;
    int mt, nt;
    tile_mn(p, tile, mt, nt);
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
    constexpr int KSW = RB / ES;   // channels per K step
    constexpr int RWB = BN * ES;   // RAWB: bytes of one weight row (one channel's BN kn)
    constexpr int CPW = RWB / 16;  // RAWB: 16-byte chunks per weight row
    constexpr int PIW = 64 / CPW;  // RAWB: weight rows per load instruction
    static_assert(!RAWB || (CPW >= (ES == 4 ? 16 : ES == 2 ? 8 : 4) && 64 % CPW == 0 && KSW / (PIW * NW) == IB &&
                            KSW % (PIW * NW) == 0), "raw weight rows");
    if constexpr (RAWB) {
        raw_rows<ES, BN, IB>(p, wave, lane, nt, b_off, b_cb, b_ok);
    } else {
#pragma unroll
        for (int j = 0; j < IB; j++) {
            const int r = (wave * IB + j) * 8 + lrow;
            const int chunk = lpos ^ ((r >> 1) & 7);
            b_cb[j] = chunk * 16;
            const int n = nt * BN + r;
            b_ok[j] = n < p.KN;
            b_off[j] = n * p.taps * CB + b_cb[j];
        }
    }

    // ---- reduction range (split-K) ----
    const int KT = p.KT;
    const int per_split = (KT + p.splits - 1) / p.splits;
    // stream-K (sk_k0 >= 0): `lin` is the tile and K steps [sk_k0, sk_k1) are reduced
    const bool skm = sk_k0 >= 0;
    const int kt_begin = skm ? sk_k0 : min(KT, split * per_split);
    const int kt_end = skm ? sk_k1 : min(KT, kt_begin + per_split);
    const int nsteps = kt_end - kt_begin;

    // issue iterator: the next K step to load (tap, channel block)
    int i_tap, i_cb;
    step_start(p, kt_begin, i_tap, i_cb);
    int i_kx = i_tap / p.KW;
    int i_ky = i_tap - i_kx * p.KW;

    auto stage_ptr = [&](auto s) __attribute__((always_inline)) -> char * {
        constexpr int S = decltype(s)::value;
        if constexpr (EXT) return ext + S * STAGE;
        else if constexpr (S == 0) return st0;
        else if constexpr (S == 1) return st1;
        else if constexpr (S == 2) return st2;
        else return st3;
    };

    // Gather one K step into a stage: LPS vm operations per wave, A rows
    // first.  issue_part sends operations [J0, J1) and advance() moves to the
    // next K step once all of them are out.  live == false loads zeros
    // (out-of-range offsets).  Issues are never skipped: a conditional issue
    // makes the compiler drain every load before the next LDS read.
    auto issue_part = [&](auto s, bool live, auto j0, auto j1) __attribute__((always_inline)) {
        constexpr int J0 = decltype(j0)::value, J1 = decltype(j1)::value;
        // The destination is passed as a plain 32-bit LDS address.  Derived
        // from the stage array, the DMA would carry that array's alias scope
        // and the compiler would add vmcnt waits of its own before later LDS
        // reads, on top of the explicit ones below; those waits cut the
        // loads in flight from NS-1 stages to one.
        const unsigned base = (unsigned)(uintptr_t)(lds_void_t *)stage_ptr(s);
        const int dh = i_kx * p.vd, dw = i_ky * p.hd;
        const int tapA = (dh * p.W + dw) * CB + i_cb * RB;
        const int lim = CB - i_cb * RB;  // channel bytes left in this block
        const int tapB = i_tap * CB + i_cb * RB;
#pragma unroll
        for (int jj = J0; jj < J1; jj++) {
            if (jj < IA) {
                const int j = jj;
                const int ih = a_ih[j] + dh, iw = a_iw[j] + dw;
                const bool ok =
                    live && (unsigned)ih < (unsigned)p.H && (unsigned)iw < (unsigned)p.W && a_cb[j] < lim;
                const unsigned off = ok ? (unsigned)(a_off[j] + tapA) : OOB;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rA, (lds_void_t *)(uintptr_t)(base + (wave * IA + j) * 1024), 16, off, 0, 0, 0);
            } else {
                const int j = jj - IA;
                bool ok;
                unsigned off;
                if constexpr (RAWB) {
                    ok = live && b_ok[j] && raw_live<ES>(p, b_cb[j], i_cb);
                    off = ok ? raw_off<ES>(p, b_off[j], i_cb, i_tap) : OOB;
                } else {
                    ok = live && b_ok[j] && b_cb[j] < lim;
                    off = ok ? (unsigned)(b_off[j] + tapB) : OOB;
                }
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rB, (lds_void_t *)(uintptr_t)(base + BM * RB + (wave * IB + j) * 1024), 16, off, 0, 0, 0);
            }
        }
    };
    auto advance = [&]() __attribute__((always_inline)) { step_next(p, i_tap, i_cb, i_kx, i_ky); };
    auto issue = [&](auto s, bool live) __attribute__((always_inline)) {
        issue_part(s, live, std::integral_constant<int, 0>{}, std::integral_constant<int, LPS>{});
        advance();
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

    // A stage holds NG groups of MFMA operands per lane: one 16-byte chunk
    // (4 fp32, used by 4 MFMAs of K = 2, or 8 bf16 for one MFMA of K = 16)
    // or two chunks (32 fp8, one MFMA of K = 64).  Lane half h takes chunk
    // 2g + h (fp8: 4g + 2h and 4g + 2h + 1) of group g for both operands, so
    // every K element of the group is summed once.  Two register sets
    // alternate: the LDS reads of group g+1 are in flight while group g
    // multiplies.
    constexpr int NG = ES == 1 ? 2 : 4;
    static_assert(NG % 2 == 0, "the last group of a stage uses register set 1");
    typedef typename std::conditional<ES == 1, i32x8, i32x4>::type OpV;
    struct Ops {
        OpV a[TR], b[TC];
    };
    // RAWB: per-lane byte offset of this lane's weight elements in each 32-column block
    int woff[TC];
    if constexpr (RAWB) raw_woff<ES, BN, TC>(lane, wn * WTN, woff);
    auto ldg = [&](auto s, int g, Ops &r) __attribute__((always_inline)) {
        const char *base = stage_ptr(s);
        if constexpr (ES != 1) {
            const int co = ((2 * g + h) ^ swz) * 16;
#pragma unroll
            for (int a = 0; a < TR; a++) r.a[a] = *(const i32x4 *)(base + rowA + a * 32 * RB + co);
            if constexpr (RAWB) {
                raw_read<ES, BN, TC>(base + BM * RB, woff, g, r.b);
            } else {
#pragma unroll
                for (int b = 0; b < TC; b++) r.b[b] = *(const i32x4 *)(base + rowB + b * 32 * RB + co);
            }
        } else {
            const int c0 = 4 * g + 2 * h;
            const int co0 = (c0 ^ swz) * 16, co1 = ((c0 + 1) ^ swz) * 16;
#pragma unroll
            for (int a = 0; a < TR; a++) {
                const i32x4 lo = *(const i32x4 *)(base + rowA + a * 32 * RB + co0);
                const i32x4 hi = *(const i32x4 *)(base + rowA + a * 32 * RB + co1);
                r.a[a] = i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
            }
            if constexpr (RAWB) {
                raw_read<ES, BN, TC>(base + BM * RB, woff, g, r.b);
            } else {
#pragma unroll
                for (int b = 0; b < TC; b++) {
                    const i32x4 lo = *(const i32x4 *)(base + rowB + b * 32 * RB + co0);
                    const i32x4 hi = *(const i32x4 *)(base + rowB + b * 32 * RB + co1);
                    r.b[b] = i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                }
            }
        }
    };
    auto mma = [&](const Ops &r) __attribute__((always_inline)) {
        if constexpr (ES == 4) {
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int a = 0; a < TR; a++)
#pragma unroll
                    for (int b = 0; b < TC; b++) {
                        // element j goes through a scalar: a bit_cast applied to
                        // a vector-element lvalue compiled to element 0 here
                        const int ea = r.a[a][j], eb = r.b[b][j];
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(__builtin_bit_cast(float, ea),
                                                                        __builtin_bit_cast(float, eb), acc[a][b], 0,
                                                                        0, 0);
                    }
        } else {
#pragma unroll
            for (int a = 0; a < TR; a++)
#pragma unroll
                for (int b = 0; b < TC; b++) {
                    if constexpr (ES == 2)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, r.a[a]),
                                                                            __builtin_bit_cast(bf16x8, r.b[b]),
                                                                            acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(r.a[a], r.b[b], acc[a][b], 0,
                                                                                    0, 0, 127, 0, 127);
                }
        }
    };

    // The operands of one stage, with the MFMAs shifted by one group: on
    // entry set 1 holds the previous stage's last group (zeros before the
    // first stage), on exit it holds this stage's last group.  Its MFMAs
    // cover the LDS latency of this stage's first reads.  Each of the NG
    // regions reads one group, multiplies the one before it and, with
    // issue_next, sends a share of the loads of step i + NS - 1, whose
    // address arithmetic fills the MFMA shadow.  The scheduling fences keep
    // the compiler from regrouping each read next to its MFMAs.
    Ops r0, r1;
#pragma unroll
    for (int a = 0; a < TR; a++) r0.a[a] = OpV{};
#pragma unroll
    for (int b = 0; b < TC; b++) r0.b[b] = OpV{};
    r1 = r0;
    constexpr int PER = (LPS + NG - 1) / NG;  // loads per region
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
        if constexpr (NG > 2) {
            region(std::integral_constant<int, (NG > 2 ? 2 : 0)>{});
            region(std::integral_constant<int, (NG > 2 ? 3 : 1)>{});
        }
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (ISS) advance();
    };
    // Before step i: this wave's loads of step i are complete (the NS-2
    // later stages may still be in flight) and so are its LDS reads of step
    // i-1; the barrier then makes every wave's part of stage i visible and
    // frees stage (i-1) % NS for the loads of step i + NS - 1.
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

    // ---- write-back (FWD NHWC: rows = output pixels, cols = kn) ----
    if (piece || p.splits > 1) {  // raw partials: stream-K piece (tile-local) or split-K slice
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int b = 0; b < TC; b++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int ml = wm * WTM + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3), nl = wn * WTN + b * 32 + l32;
                    if (piece) {
                        piece[ml * BN + nl] = acc[a][b][v];
                        continue;
                    }
                    const int m = mt * BM + ml, n = nt * BN + nl;
                    if (m < x.Mrows && n < p.KN) p.ws[((long long)split * p.M + m) * p.N + n] = acc[a][b][v];
                }
        return;
    }
    if (!epi_staged<CG_FWD_NHWC>(p) && p.beta == 0.0f) {  // plain: nothing to load, one store per output
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int b = 0; b < TC; b++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int ml = wm * WTM + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3), nl = wn * WTN + b * 32 + l32;
                    const int m = mt * BM + ml, n = nt * BN + nl;
                    if (m < x.Mrows && n < p.KN) store_out<CG_FWD_NHWC, (ES < 4)>(p, x, m, n, acc[a][b][v]);
                }
        return;
    }
    // the fused epilogue: per-channel vectors staged in LDS (the pipeline stages are dead)
    float *s_epi = nullptr;
    if (epi_staged<CG_FWD_NHWC>(p)) {
        static_assert(5 * BN * sizeof(float) <= STAGE, "epilogue vectors fit one LDS stage");
        __syncthreads();
        s_epi = reinterpret_cast<float *>(stage_ptr(std::integral_constant<int, 0>{}));
        epi_stage(p, nt * BN, BN, s_epi, tid, NW * 64);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int b = 0; b < TC; b++)
            store_block<CG_FWD_NHWC, (ES < 4)>(p, x, acc[a][b], s_epi, BN, [&](int v, int &m, int &n, int &nl) __attribute__((always_inline)) {
                const int ml = wm * WTM + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
                nl = wn * WTN + b * 32 + l32;
                m = mt * BM + ml;
                n = nt * BN + nl;
                return m < x.Mrows && n < p.KN;
            });
    if (s_epi) __syncthreads();  // the LDS is rewritten by the next item of a persistent launch
}

template <typename T, int BM, int BN, int WM, int WN, int NS, bool RAWB>
__global__ void __launch_bounds__(WM * WN * 64) lp_kernel(const cg_kparams p) {
    lp_tile<T, BM, BN, WM, WN, NS, false, RAWB>(p, xcd_run(blockIdx.x, gridDim.x));
}

// Stream-K: blocks 0 .. sk_dp-1 each run one whole tile of the first sk_dp
// (data-parallel); the K steps of the remaining tiles are cut into p.sk equal
// contiguous ranges, one per later block.  A whole tile is written back
// directly; a partial one (only the first and last of a range can be) goes
// to the workgroup's piece 0 / 1, and sk_fixup_kernel sums the pieces of
// every shared tile in workgroup order.  sk_dp is a multiple of 8, so the XCD
// of a block of either region is its index within the region mod 8.
//
// Both regions run one loop with a single call of the tile body, and the
// parameters are read through the kernarg pointer where used: with a second
// inlined copy of the body and the parameters held as values from the kernel
// entry, the scalar state live across the body overflowed the SGPRs and the
// spills were reloaded with v_readlane all through the K loop (2580 reloads
// in the 64x128 NCHW kernel, against 21 now).
template <typename T, int BM, int BN, int WM, int WN, int NS, bool RAWB>
__global__ void __launch_bounds__(WM * WN * 64) lp_sk_kernel(const cg_kparams p_arg) {
    const cg_kparams &p = kernarg_params(p_arg);
    SkRange r = sk_range(p);
    for (SkPiece q; sk_next(p, r, q, BM * BN);) {
        lp_tile<T, BM, BN, WM, WN, NS, false, RAWB>(p, q.tile, q.k0, q.k1, q.piece);
        __syncthreads();  // loads, LDS reads and stores of this piece are complete
    }
}

// Grouped form, with the item layout of igemm_grouped_kernel.
template <typename T, int BM, int BN, int WM, int WN, int NS>
__global__ void __launch_bounds__(WM * WN * 64) lp_grouped_kernel(const cg_kparams *__restrict__ plist,
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
        if (lin < count) lp_tile<T, BM, BN, WM, WN, NS>(p, lin);
    }
}

// Forward NHWC with 16-byte channel chunks only.  kp->braw: the weights are
// the caller's [c][kh][kw][kn] (plain and stream-K launches).
template <typename T, int BM, int BN, int WM, int WN, int NS>
inline int launch_lp(const cg_kparams *kp, const cg_group *g, int *occ, hipStream_t st) {
    constexpr int NT = WM * WN * 64;
    // FP8 wave tiles of 128x64 and up hold 128 accumulator registers beside
    // 32-byte operands: their stream-K form spills, and is not built
    constexpr bool kSK = !(sizeof(T) == 1 && (BM / WM) * (BN / WN) >= 128 * 64);
    if (kp->op != CG_FWD_NHWC || !kp->vec) return (int)hipErrorInvalidValue;
    if (occ) {
        if (kp->sk < 0) {
            if constexpr (!kSK) {
                *occ = 0;
                return 0;
            } else {
                return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    occ, reinterpret_cast<const void *>(&lp_sk_kernel<T, BM, BN, WM, WN, NS, false>), NT, 0);
            }
        }
        return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            occ, reinterpret_cast<const void *>(&lp_grouped_kernel<T, BM, BN, WM, WN, NS>), NT, 0);
    }
    if (kp->braw && g) return (int)hipErrorInvalidValue;  // grouped launches read repacked weights
    if (kp->sk > 0 && !g) {
        if constexpr (!kSK) {
            return (int)hipErrorInvalidValue;
        } else {
            const dim3 grid((unsigned)(kp->sk_dp + kp->sk));
            if (kp->braw) {
                hipLaunchKernelGGL((lp_sk_kernel<T, BM, BN, WM, WN, NS, true>), grid, dim3(NT), 0, st, *kp);
                return (int)hipGetLastError();
            }
            hipLaunchKernelGGL((lp_sk_kernel<T, BM, BN, WM, WN, NS, false>), grid, dim3(NT), 0, st, *kp);
            return (int)hipGetLastError();
        }
    }
    if (g) {
        if (g->total <= 0) return 0;
        hipLaunchKernelGGL((lp_grouped_kernel<T, BM, BN, WM, WN, NS>), dim3((unsigned)g->grid), dim3(NT), 0, st,
                           g->plist, g->start, g->nprob, g->total);
        return (int)hipGetLastError();
    }
    const long long nb = (long long)kp->mtiles * kp->ntiles * kp->splits;
    if (kp->braw) {
        hipLaunchKernelGGL((lp_kernel<T, BM, BN, WM, WN, NS, true>), dim3((unsigned)nb), dim3(NT), 0, st, *kp);
        return (int)hipGetLastError();
    }
    hipLaunchKernelGGL((lp_kernel<T, BM, BN, WM, WN, NS, false>), dim3((unsigned)nb), dim3(NT), 0, st, *kp);
    return (int)hipGetLastError();
}

}  // namespace cg
