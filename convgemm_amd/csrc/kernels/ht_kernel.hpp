// ht_kernel.hpp -- halo-tile implicit-GEMM forward kernel (NHWC; FP32, BF16, FP8).
//
// The forward NHWC convolution of src/convGemm.c:48-70 for kernels larger than
// 1x1.  The pipelined kernel (lp_kernel.hpp) gathers the patch rows of every
// tap from HBM/L2 again: a 3x3 stride-1 layer loads each input pixel about 9
// times per tile, and the load path, not the matrix cores, sets its time.
// Here an M tile is whole output rows of ONE image (hrows rows, when a row
// fits the BM tile rows), or one of hcblk pieces of a row, and for each
// channel chunk (128 bytes of channels) the input pixels those outputs need --
// the halo, (rows - 1) * stride + (KH - 1) * dilation + 1 input rows of
// HC = (cols - 1) * hstride + (KW - 1) * hdilation + 1 slots, a slot per input
// column from the tile's first one on, zero where it is padding -- is loaded
// into LDS once.  Tiles of whole rows keep the halo free of the partial rows
// a run of BM pixels would straddle (the tile rows past rows * Wo idle).  The KH*KW taps of the chunk then read their A operands from that
// buffer at tap-dependent offsets; only the weight slices (BN rows of 128
// bytes per tap) stream through NS LDS stages as in the pipelined kernel.
//
//   K steps run chunk-major (all taps of a chunk, then the next chunk); split-K
//   divides the chunks.  The halo buffer is single: at a chunk boundary the
//   next halo is loaded after every wave has read the last one (its latency is
//   covered by the other workgroup resident on the CU).
//   LDS rows are 128 bytes with the 16-byte chunk c of slot s at position
//   c ^ ((s >> 1) & 7) (the pipelined kernel's swizzle, for slots).
//   The host picks a configuration only when the largest halo of the problem
//   fits its HS slots (cg_halo_fits).
#pragma once
#include "lp_kernel.hpp"

namespace cg {

// RAWB: p.Bm holds the caller's [c][kh][kw][kn] weights, read in place as in
// the pipelined kernel (raw_rows / raw_read): a weight stage is one row per
// channel of the chunk, holding the tile's BN kn.
template <typename T, int BM, int BN, int WM, int WN, int NS, int HS, bool RAWB = false>
__device__ __attribute__((always_inline)) inline void ht_tile(const cg_kparams &p, const int lin) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 2 || sizeof(T) == 1, "FP32, BF16 or FP8 operands");
    constexpr int ES = sizeof(T);
    constexpr int RB = 128;                 // bytes per row (slot) per chunk
    constexpr int NW = WM * WN;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int TR = WTM / 32, TC = WTN / 32;
    constexpr int IB = BN / (8 * NW);       // weight load instructions per wave per step
    constexpr int NI = HS / (8 * NW);       // halo load instructions per wave per chunk
    constexpr int WST = BN * RB;            // bytes of one weight stage
    static_assert((NW == 4 || NW == 8) && TR >= 1 && TC >= 1 && WTM % 32 == 0 && WTN % 32 == 0, "wave tile");
    static_assert(BN % (8 * NW) == 0 && HS % (8 * NW) == 0 && NS >= 2 && NS <= 4, "tile");
    static_assert(NI + (NS - 1) * IB < 64, "vmcnt");
    constexpr int CPW = BN * ES / 16;  // RAWB: 16-byte chunks per weight row
    static_assert(!RAWB || (CPW >= (ES == 4 ? 16 : ES == 2 ? 8 : 4) && 64 % CPW == 0 &&
                            (RB / ES) / ((64 / CPW) * NW) == IB && (RB / ES) % ((64 / CPW) * NW) == 0),
                  "raw weight rows");

    __shared__ __attribute__((aligned(1024))) char sh[HS * RB];
    __shared__ __attribute__((aligned(1024))) char sw0[WST];
    __shared__ __attribute__((aligned(1024))) char sw1[WST];
    __shared__ __attribute__((aligned(1024))) char sw2[NS > 2 ? WST : 16];
    __shared__ __attribute__((aligned(1024))) char sw3[NS > 3 ? WST : 16];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: LDS addresses go to M0

    const int tiles = p.mtiles * p.ntiles;
    const int tile = lin % tiles, split = lin / tiles;
    int mt, nt;
    tile_mn(p, tile, mt, nt);
    const Ctx x = make_ctx(p, 0);

    // ---- the tile's pixels and its halo ----
    const int HoWo = p.Ho * p.Wo;
    const int CBK = p.hcblk;
    const int b = mt / p.htpi;
    const int tin = mt - b * p.htpi;  // tile within the image
    int trow = tin, tcol = 0;
    if (CBK > 1) {  // pieces of a row (uniform)
        trow = tin / CBK;
        tcol = tin - trow * CBK;
    }
    const int r_first = trow * p.hrows, c_first = tcol * BM;  // c_first is 0 for tiles of whole rows
    const int nrows = min(p.hrows, p.Ho - r_first);
    const int ncols = CBK == 1 ? p.Wo : min(BM, p.Wo - c_first);
    const int count = nrows * ncols;        // output pixels of the tile: q0 .. q0 + count - 1 of image b
    const int q0 = r_first * p.Wo + c_first;
    const int ih0 = r_first * p.vs - p.vp;  // input row of halo row 0
    const int iw0 = c_first * p.hs - p.hp;  // input column of halo column 0
    // halo rows are HC slots apart, HC that of a full tile (a partial tile loads columns it never reads)
    const int HC = p.hpitch;
    const int HR = (nrows - 1) * p.vs + (p.KH - 1) * p.vd + 1;
    const int slots = HR * HC;
    const FastDiv fhc = host_fastdiv(p.fdm_hp, p.fds_hp, HC);

    const __amdgpu_buffer_rsrc_t rA = make_rsrc(p.A, p.a_bytes);
    const __amdgpu_buffer_rsrc_t rB = make_rsrc(p.Bm, p.b_bytes);
    const int CB = p.C * ES;  // bytes of one pixel's channels (a multiple of 16)

    // ---- halo loads: lane l of instruction k of this wave fills slot
    // (wave*NI + k)*8 + l/8, position l%8, with global chunk pos ^ swz(slot) ----
    const int lrow = lane >> 3, lpos = lane & 7;
    unsigned h_off[NI];
    int h_cb[NI];
#pragma unroll
    for (int k = 0; k < NI; k++) {
        const int s = (wave * NI + k) * 8 + lrow;
        const int chunk = lpos ^ ((s >> 1) & 7);
        h_cb[k] = chunk * 16;
        const int hr = fdiv(s, fhc), hc = s - hr * HC;
        const int ih = ih0 + hr, iw = iw0 + hc;
        const bool ok = s < slots && (unsigned)ih < (unsigned)p.H && (unsigned)iw < (unsigned)p.W;
        h_off[k] = ok ? (unsigned)(((b * p.H + ih) * p.W + iw) * CB + h_cb[k]) : OOB;
    }
    // ---- weight rows ([kn][tap][c] repack, or in place): as in the pipelined kernel ----
    int b_off[IB], b_cb[IB];
    bool b_ok[IB];
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

    // ---- reduction range: split-K divides the channel chunks ----
    const int nch = p.kchunks;
    const int per_split = (nch + p.splits - 1) / p.splits;
    const int c_begin = min(nch, split * per_split), c_end = min(nch, c_begin + per_split);
    const int T_ = p.taps;
    const int nsteps = (c_end - c_begin) * T_;

    // issue iterator of the weight stream: the next step's (chunk, tap)
    int i_cb = c_begin, i_tap = 0;

    auto wstage = [&](auto s) __attribute__((always_inline)) -> char * {
        constexpr int S = decltype(s)::value;
        if constexpr (S == 0) return sw0;
        else if constexpr (S == 1) return sw1;
        else if constexpr (S == 2) return sw2;
        else return sw3;
    };
    // Weight loads [J0, J1) of the next step into stage s (live == false loads
    // zeros: an issue is never skipped, so the compiler adds no waits).
    auto issue_part = [&](auto s, bool live, auto j0, auto j1) __attribute__((always_inline)) {
        constexpr int J0 = decltype(j0)::value, J1 = decltype(j1)::value;
        const unsigned base = (unsigned)(uintptr_t)(lds_void_t *)wstage(s);
        const int lim = CB - i_cb * RB;
        const int tapB = i_tap * CB + i_cb * RB;
#pragma unroll
        for (int j = J0; j < J1; j++) {
            bool ok;
            unsigned off;
            if constexpr (RAWB) {
                ok = live && b_ok[j] && raw_live<ES>(p, b_cb[j], i_cb);
                off = ok ? raw_off<ES>(p, b_off[j], i_cb, i_tap) : OOB;
            } else {
                ok = live && b_ok[j] && b_cb[j] < lim;
                off = ok ? (unsigned)(b_off[j] + tapB) : OOB;
            }
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rB, (lds_void_t *)(uintptr_t)(base + (wave * IB + j) * 1024), 16,
                                                     off, 0, 0, 0);
        }
    };
    auto advance = [&]() __attribute__((always_inline)) {
        if (++i_tap == T_) {
            i_tap = 0;
            ++i_cb;
        }
    };
    auto issue = [&](auto s, bool live) __attribute__((always_inline)) {
        issue_part(s, live, std::integral_constant<int, 0>{}, std::integral_constant<int, IB>{});
        advance();
    };
    // The halo of channel chunk cb: NI loads per wave.
    auto issue_halo = [&](int cb) __attribute__((always_inline)) {
        const unsigned base = (unsigned)(uintptr_t)(lds_void_t *)sh;
        const int lim = CB - cb * RB;
#pragma unroll
        for (int k = 0; k < NI; k++) {
            const unsigned off = (h_off[k] != OOB && h_cb[k] < lim) ? h_off[k] + (unsigned)(cb * RB) : OOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rA, (lds_void_t *)(uintptr_t)(base + (wave * NI + k) * 1024), 16,
                                                     off, 0, 0, 0);
        }
    };

    // ---- MFMA: A from the halo (per-lane slot + the tap's offset), B from a weight stage ----
    const int wm = wave % WM, wn = wave / WM;
    const int l32 = lane & 31, h = lane >> 5;
    const int swz_b = (l32 >> 1) & 7;  // weight rows of this lane are l32 mod 32
    const int rowB = (wn * WTN + l32) * RB;
    int woff[TC];  // RAWB: this lane's offsets in a weight stage
    if constexpr (RAWB) raw_woff<ES, BN, TC>(lane, wn * WTN, woff);
    int aslot[TR];  // halo slot of this lane's output pixel in each 32-row block (tap 0)
#pragma unroll
    for (int a = 0; a < TR; a++) {
        const int m = wm * WTM + a * 32 + l32;
        aslot[a] = 0;  // rows past the tile read slot 0 and are not stored
        if (m < count) {
            const int orow = CBK == 1 ? fdiv(m, x.wo) : 0, ocol = m - orow * ncols;  // ncols == Wo for whole rows
            aslot[a] = orow * p.vs * HC + ocol * p.hs;
        }
    }
    f32x16 acc[TR][TC];
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int c = 0; c < TC; c++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[a][c][v] = 0.f;

    constexpr int NG = ES == 1 ? 2 : 4;
    typedef typename std::conditional<ES == 1, i32x8, i32x4>::type OpV;
    struct Ops {
        OpV a[TR], b[TC];
    };
    auto ldg = [&](auto s, int g, int tapoff, Ops &r) __attribute__((always_inline)) {
        const char *wb = wstage(s);
        if constexpr (ES != 1) {
#pragma unroll
            for (int a = 0; a < TR; a++) {
                const int sl = aslot[a] + tapoff;
                r.a[a] = *(const i32x4 *)(sh + sl * RB + (((2 * g + h) ^ ((sl >> 1) & 7)) * 16));
            }
            const int co = ((2 * g + h) ^ swz_b) * 16;
            if constexpr (RAWB) {
                raw_read<ES, BN, TC>(wb, woff, g, r.b);
            } else {
#pragma unroll
                for (int c = 0; c < TC; c++) r.b[c] = *(const i32x4 *)(wb + rowB + c * 32 * RB + co);
            }
        } else {
            const int c0 = 4 * g + 2 * h;
#pragma unroll
            for (int a = 0; a < TR; a++) {
                const int sl = aslot[a] + tapoff;
                const int sz = (sl >> 1) & 7;
                const i32x4 lo = *(const i32x4 *)(sh + sl * RB + ((c0 ^ sz) * 16));
                const i32x4 hi = *(const i32x4 *)(sh + sl * RB + (((c0 + 1) ^ sz) * 16));
                r.a[a] = i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
            }
            const int co0 = (c0 ^ swz_b) * 16, co1 = ((c0 + 1) ^ swz_b) * 16;
            if constexpr (RAWB) {
                raw_read<ES, BN, TC>(wb, woff, g, r.b);
            } else {
#pragma unroll
                for (int c = 0; c < TC; c++) {
                    const i32x4 lo = *(const i32x4 *)(wb + rowB + c * 32 * RB + co0);
                    const i32x4 hi = *(const i32x4 *)(wb + rowB + c * 32 * RB + co1);
                    r.b[c] = i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
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
                    for (int c = 0; c < TC; c++) {
                        const int ea = r.a[a][j], eb = r.b[c][j];  // through scalars (see lp_kernel.hpp)
                        acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(__builtin_bit_cast(float, ea),
                                                                        __builtin_bit_cast(float, eb), acc[a][c], 0,
                                                                        0, 0);
                    }
        } else {
#pragma unroll
            for (int a = 0; a < TR; a++)
#pragma unroll
                for (int c = 0; c < TC; c++) {
                    if constexpr (ES == 2)
                        acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, r.a[a]),
                                                                            __builtin_bit_cast(bf16x8, r.b[c]),
                                                                            acc[a][c], 0, 0, 0);
                    else
                        acc[a][c] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(r.a[a], r.b[c], acc[a][c], 0,
                                                                                    0, 0, 127, 0, 127);
                }
        }
    };

    Ops r0, r1;
#pragma unroll
    for (int a = 0; a < TR; a++) r0.a[a] = OpV{};
#pragma unroll
    for (int c = 0; c < TC; c++) r0.b[c] = OpV{};
    r1 = r0;
    constexpr int PER = (IB + NG - 1) / NG;  // weight loads per region
    // One step as in the pipelined kernel: NG fenced regions, each reading the
    // operands of one group, multiplying the previous one and sending a share
    // of the weight loads of step i + NS - 1.  Set 1 carries the last group of
    // a step into the next.
    // compute-side tap state of the current step: tap index, column of the tap
    // in the kernel, and its halo offset kx * vd * HC + ky * hd
    int c_tap = 0, c_ky = 0, c_off = 0, c_ch = 0;  // c_ch: channel chunk of the step, from c_begin
    const int row_step = p.vd * HC - (p.KW - 1) * p.hd;  // offset change from (kx, KW - 1) to (kx + 1, 0)
    auto next_tap = [&]() __attribute__((always_inline)) {  // selects, no branches in the step
        const bool wrap_tap = c_tap + 1 == T_, wrap_ky = c_ky + 1 == p.KW;
        c_off = wrap_tap ? 0 : c_off + (wrap_ky ? row_step : p.hd);
        c_ky = wrap_tap || wrap_ky ? 0 : c_ky + 1;
        c_tap = wrap_tap ? 0 : c_tap + 1;
        c_ch += wrap_tap ? 1 : 0;
    };
    auto stage_mma = [&](auto s, auto issue_next, int i) __attribute__((always_inline)) {
        constexpr bool ISS = decltype(issue_next)::value;
        constexpr int SN = (decltype(s)::value + NS - 1) % NS;
        const bool live = i + NS - 1 < nsteps;
        const int tapoff = c_off;
        auto region = [&](auto g) __attribute__((always_inline)) {
            constexpr int G = decltype(g)::value;
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (G % 2 == 0) {
                ldg(s, G, tapoff, r0);
                mma(r1);
            } else {
                ldg(s, G, tapoff, r1);
                mma(r0);
            }
            if constexpr (ISS && G * PER < IB)
                issue_part(std::integral_constant<int, SN>{}, live, std::integral_constant<int, G * PER>{},
                           std::integral_constant<int, (G * PER + PER < IB ? G * PER + PER : IB)>{});
        };
        region(std::integral_constant<int, 0>{});
        region(std::integral_constant<int, 1>{});
        if constexpr (NG > 2) {
            region(std::integral_constant<int, (NG > 2 ? 2 : 0)>{});
            region(std::integral_constant<int, (NG > 2 ? 3 : 1)>{});
        }
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (ISS) advance();
        next_tap();
    };
    // Before step i: the weight loads of step i have landed (those of the NS-2
    // later steps may be in flight) and every wave's LDS reads of step i-1 are
    // done; the barrier publishes stage i and frees stage (i-1) % NS.  At the
    // first step of a chunk after the first, the halo of that chunk is loaded
    // over the previous one (all its readers are past the barrier) and waited
    // for in full.
    auto sync = [&](auto pending, int i) __attribute__((always_inline)) {
        wait_vm<decltype(pending)::value>();
        __builtin_amdgcn_s_waitcnt((0xF) | (0x7 << 4) | (0x0 << 8) | (0x3 << 14));  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        if (i > 0 && c_tap == 0) {  // first step of a chunk (uniform)
            issue_halo(c_begin + c_ch);
            wait_vm<0>();
            __builtin_amdgcn_s_barrier();
        }
    };
    constexpr auto kInFlight = std::integral_constant<int, (NS - 2) * IB>{};
    constexpr auto kYes = std::true_type{};
    constexpr auto kNo = std::false_type{};

    if (nsteps > 0) {
        issue_halo(c_begin);
        issue(std::integral_constant<int, 0>{}, true);
        if constexpr (NS > 2) issue(std::integral_constant<int, 1>{}, 1 < nsteps);
        if constexpr (NS > 3) issue(std::integral_constant<int, (NS > 3 ? 2 : 0)>{}, 2 < nsteps);
        int i = 0;
        for (; i + NS <= nsteps; i += NS) {
            sync(kInFlight, i);
            stage_mma(std::integral_constant<int, 0>{}, kYes, i);
            sync(kInFlight, i + 1);
            stage_mma(std::integral_constant<int, 1>{}, kYes, i + 1);
            if constexpr (NS > 2) {
                sync(kInFlight, i + 2);
                stage_mma(std::integral_constant<int, (NS > 2 ? 2 : 0)>{}, kYes, i + 2);
            }
            if constexpr (NS > 3) {
                sync(kInFlight, i + 3);
                stage_mma(std::integral_constant<int, (NS > 3 ? 3 : 0)>{}, kYes, i + 3);
            }
        }
        // tail: the last nsteps % NS steps sit in stages 0.. and are all issued
        const int rem = nsteps - i;
        if (rem > 0) {
            sync(std::integral_constant<int, 0>{}, i);
            stage_mma(std::integral_constant<int, 0>{}, kNo, i);
            if (rem > 1) {
                sync(std::integral_constant<int, 0>{}, i + 1);
                stage_mma(std::integral_constant<int, 1>{}, kNo, i + 1);
            }
            if constexpr (NS > 3)
                if (rem > 2) {
                    sync(std::integral_constant<int, 0>{}, i + 2);
                    stage_mma(std::integral_constant<int, (NS > 3 ? 2 : 0)>{}, kNo, i + 2);
                }
        }
        mma(r1);
    }

    // ---- write-back: rows = output pixels q0 .. q0 + count - 1 of image b, cols = kn ----
    const int pbase = b * HoWo + q0;
    if (p.splits > 1) {  // raw split-K partials
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int c = 0; c < TC; c++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int ml = wm * WTM + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3), nl = wn * WTN + c * 32 + l32;
                    const int n = nt * BN + nl;
                    if (ml < count && n < p.KN)
                        p.ws[((long long)split * p.M + pbase + ml) * p.N + n] = acc[a][c][v];
                }
        return;
    }
    if (!epi_staged<CG_FWD_NHWC>(p) && p.beta == 0.0f) {  // plain: nothing to load, one store per output
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
            for (int c = 0; c < TC; c++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int ml = wm * WTM + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3), nl = wn * WTN + c * 32 + l32;
                    const int n = nt * BN + nl;
                    if (ml < count && n < p.KN) store_out<CG_FWD_NHWC, (sizeof(T) < 4)>(p, x, pbase + ml, n, acc[a][c][v]);
                }
        return;
    }
    // the fused epilogue: per-channel vectors staged in LDS (the halo is dead)
    float *s_epi = nullptr;
    if (epi_staged<CG_FWD_NHWC>(p)) {
        static_assert(5 * BN * sizeof(float) <= HS * RB, "epilogue vectors fit the halo buffer");
        __syncthreads();
        s_epi = reinterpret_cast<float *>(sh);
        epi_stage(p, nt * BN, BN, s_epi, tid, NW * 64);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int c = 0; c < TC; c++)
            store_block<CG_FWD_NHWC, (sizeof(T) < 4)>(p, x, acc[a][c], s_epi, BN,
                                     [&](int v, int &m, int &n, int &nl) __attribute__((always_inline)) {
                                         const int ml = wm * WTM + a * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
                                         nl = wn * WTN + c * 32 + l32;
                                         m = pbase + ml;
                                         n = nt * BN + nl;
                                         return ml < count && n < p.KN;
                                     });
    if (s_epi) __syncthreads();  // the LDS is rewritten by the next item of a persistent launch
}

template <typename T, int BM, int BN, int WM, int WN, int NS, int HS, bool RAWB>
__global__ void __launch_bounds__(WM * WN * 64) ht_kernel(const cg_kparams p) {
    ht_tile<T, BM, BN, WM, WN, NS, HS, RAWB>(p, xcd_run(blockIdx.x, gridDim.x));
}

// Grouped form, with the item layout of igemm_grouped_kernel.
template <typename T, int BM, int BN, int WM, int WN, int NS, int HS>
__global__ void __launch_bounds__(WM * WN * 64) ht_grouped_kernel(const cg_kparams *__restrict__ plist,
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
        if (lin < count) ht_tile<T, BM, BN, WM, WN, NS, HS>(p, lin);
    }
}

// Forward NHWC with 16-byte channel chunks; no stream-K form.  kp->braw: the
// weights are the caller's [c][kh][kw][kn] (plain launches).
template <typename T, int BM, int BN, int WM, int WN, int NS, int HS>
inline int launch_ht(const cg_kparams *kp, const cg_group *g, int *occ, hipStream_t st) {
    constexpr int NT = WM * WN * 64;
    if (kp->op != CG_FWD_NHWC || !kp->vec) return (int)hipErrorInvalidValue;
    if (occ) {
        if (kp->sk < 0) {  // the stream-K kernel's occupancy: there is none
            *occ = 0;
            return 0;
        }
        return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            occ, reinterpret_cast<const void *>(&ht_grouped_kernel<T, BM, BN, WM, WN, NS, HS>), NT, 0);
    }
    if (kp->sk > 0 || (kp->braw && g)) return (int)hipErrorInvalidValue;
    if (g) {
        if (g->total <= 0) return 0;
        hipLaunchKernelGGL((ht_grouped_kernel<T, BM, BN, WM, WN, NS, HS>), dim3((unsigned)g->grid), dim3(NT), 0, st,
                           g->plist, g->start, g->nprob, g->total);
        return (int)hipGetLastError();
    }
    const long long nb = (long long)kp->mtiles * kp->ntiles * kp->splits;
    if (kp->braw) {
        hipLaunchKernelGGL((ht_kernel<T, BM, BN, WM, WN, NS, HS, true>), dim3((unsigned)nb), dim3(NT), 0, st, *kp);
        return (int)hipGetLastError();
    }
    hipLaunchKernelGGL((ht_kernel<T, BM, BN, WM, WN, NS, HS, false>), dim3((unsigned)nb), dim3(NT), 0, st, *kp);
    return (int)hipGetLastError();
}

}  // namespace cg
