// igemm_aux.hip -- kernels around the implicit GEMM and the launcher entry
// points of cg_internal.h:
//   split-K reduction (deterministic, fixed summation order) + epilogue,
//   beta scaling (the alpha == 0 path of src/quick_gemm.h),
//   weight repacks (LDS-tiled strided transpose, 4/2/1-byte elements),
//   FP32 -> BF16 / FP8-e4m3 conversion with per-tensor amax scaling,
//   per-element-type dispatch of the implicit-GEMM configurations.
#include "igemm_common.hpp"

namespace cg {

// ---- split-K reduction + epilogue ----
// A 256-thread block reduces 256 / G consecutive outputs of one class, with G
// split groups per output: group g sums the splits g, g + G, g + 2G, ... and
// group 0 adds the other groups' sums in order (a fixed summation order, the
// same in the single and the grouped launch).  Deep splits over small outputs
// (the weight gradient: fewer than 512 blocks of 256 outputs) get G = 4: four
// times the blocks and a quarter of the serial chain of one thread per output.
// Large outputs keep G = 1 (there the extra blocks' set-up cost more).
__host__ __device__ inline int reduce_groups(const cg_kparams &p) {
    return p.splits >= 8 && (long long)p.M * p.N < 512LL * 256 ? 4 : 1;
}

template <int OP>
__device__ __attribute__((always_inline)) inline void reduce_block(const cg_kparams &p, int cls, long long blk) {
    __shared__ float part[3][64];
    const int G = reduce_groups(p);  // uniform over the block
    const int per = 256 / G;
    const int g = (int)threadIdx.x / per, e = (int)threadIdx.x - g * per;
    const Ctx x = make_ctx(p, cls);
    const long long total = (long long)x.Mrows * p.N;
    const long long idx = blk * per + e;
    const bool live = idx < total;
    const int m = live ? (int)(idx / p.N) : 0, n = live ? (int)(idx - (long long)m * p.N) : 0;
    const int ncls = p.ncls_h * p.ncls_w;
    float s = 0.f;
    if (live)
        for (int z = g; z < p.splits; z += G) s += p.ws[(((long long)z * ncls + cls) * p.M + m) * p.N + n];
    if (G == 1) {
        if (live) store_out<OP>(p, x, m, n, s);
        return;
    }
    if (g > 0) part[g - 1][e] = s;
    __syncthreads();
    if (g == 0 && live) {
        s += part[0][e];
        s += part[1][e];
        s += part[2][e];
        store_out<OP>(p, x, m, n, s);
    }
    __syncthreads();  // `part` is reused by the next item of a grouped launch
}

template <int OP>
__global__ void __launch_bounds__(256) splitk_reduce_kernel(const cg_kparams p) {
    reduce_block<OP>(p, blockIdx.y, blockIdx.x);
}

// Grouped: item = one block of one class of one problem; the blocks of a
// class cover M*N (the largest class), as in the single launch.
template <int OP>
__global__ void __launch_bounds__(256) splitk_reduce_grouped_kernel(const cg_kparams *__restrict__ plist,
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
        const int local = item - start[lo];
        const int per = 256 / reduce_groups(p);
        const int per_cls = (int)(((long long)p.M * p.N + per - 1) / per);
        const int cls = local / per_cls, blk = local - cls * per_cls;
        reduce_block<OP>(p, cls, blk);
    }
}

// Grouped over problems of any operations: the op of each item's problem
// picks the instantiation (uniform per item).
__global__ void __launch_bounds__(256) splitk_reduce_mixed_kernel(const cg_kparams *__restrict__ plist,
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
        const int local = item - start[lo];
        const int per = 256 / reduce_groups(p);
        const int per_cls = (int)(((long long)p.M * p.N + per - 1) / per);
        const int cls = local / per_cls, blk = local - cls * per_cls;
        switch (p.op) {
        case CG_FWD_NHWC: reduce_block<CG_FWD_NHWC>(p, cls, blk); break;
        case CG_FWD_NCHW: reduce_block<CG_FWD_NCHW>(p, cls, blk); break;
        case CG_WGRAD_NHWC: reduce_block<CG_WGRAD_NHWC>(p, cls, blk); break;
        case CG_WGRAD_NCHW: reduce_block<CG_WGRAD_NCHW>(p, cls, blk); break;
        case CG_DGRAD_NHWC: reduce_block<CG_DGRAD_NHWC>(p, cls, blk); break;
        default: reduce_block<CG_DGRAD_NCHW>(p, cls, blk); break;
        }
    }
}

// ---- stream-K fix-up: one workgroup per tile of the stream-K region (tiles
// sk_dp ..); tiles that several stream-K workgroups share sum their pieces
// (workgroup order, deterministic) and go through the epilogue; whole tiles
// were written back by their workgroup.
//
// A tile has BM x BN = 2^lbm x 2^lbn elements (every tile of the tables is
// at least 64 x 64, a multiple of 2048).  Pieces are [BM][BN] floats, except
// for the NCHW forward pass ([BN][BM]: pixels contiguous, as the output).
// Block b takes elements [2048 c, 2048 c + 2048) of stream-K tile b / (TE /
// 2048), c = b mod (TE / 2048): each thread sums 8 elements over all pieces,
// then (beta != 0) loads their old values, then stores the 8, so no store
// sits between loads that could overlap.  The per-channel epilogue vectors
// of the tile's columns are staged in LDS. ----
constexpr int FIX_ELEMS = 8 * 256;  // elements per fix-up block

template <int OP>
__global__ void __launch_bounds__(256) sk_fixup_kernel(const cg_kparams p, int lbm, int lbn) {
    constexpr bool WG = OP == CG_WGRAD_NHWC || OP == CG_WGRAD_NCHW;
    constexpr bool FW = OP == CG_FWD_NHWC || OP == CG_FWD_NCHW;
    constexpr bool PM = OP == CG_FWD_NCHW;  // pixel-major pieces
    __shared__ float s_epi[5 * 256];
    const int BM = 1 << lbm, BN = 1 << lbn, TE = BM * BN;
    const int tid = threadIdx.x;
    const int per_tile = TE / FIX_ELEMS;
    const int tl = (int)blockIdx.x / per_tile;  // stream-K tile of this block
    const int base = ((int)blockIdx.x - tl * per_tile) * FIX_ELEMS + tid;
    const int t = p.sk_dp + tl;
    const long long G = p.sk;
    const long long U = ((long long)p.mtiles * p.ntiles - p.sk_dp) * p.KT;
    const long long lo = (long long)tl * p.KT, hi = lo + p.KT;
    auto first_unit = [&](long long g) { return g * U / G; };
    long long g0 = lo * G / U;
    while (g0 + 1 < G && first_unit(g0 + 1) <= lo) g0++;
    long long g1 = (hi - 1) * G / U;
    while (g1 + 1 < G && first_unit(g1 + 1) <= hi - 1) g1++;
    if (g0 == g1) return;  // uniform: one workgroup wrote the whole tile
    const Ctx x = make_ctx(p, 0);
    int mt, nt;
    tile_mn(p, t, mt, nt);
    int wg_tap = 0, wg_c0 = 0;
    if (WG) {
        wg_tap = mt / p.kchunks;
        wg_c0 = (mt - wg_tap * p.kchunks) * BM;
    }
    const bool staged = epi_staged<OP>(p) && BN <= 256;
    if (staged) {
        epi_stage(p, nt * BN, BN, s_epi, tid, 256);
        __syncthreads();
    }
    auto coords = [&](int e, int &m, int &n, int &nl) __attribute__((always_inline)) {
        int ml;
        if (PM) {
            nl = e >> lbm;
            ml = e & (BM - 1);
        } else {
            ml = e >> lbn;
            nl = e & (BN - 1);
        }
        n = nt * BN + nl;
        if (WG) {
            const int c = wg_c0 + ml;
            m = wg_tap * p.C + c;
            return c < p.C && n < p.KN;
        }
        m = mt * BM + ml;
        return m < x.Mrows && n < p.KN;
    };
    {
        float acc[8];
#pragma unroll
        for (int j = 0; j < 8; j++) acc[j] = 0.f;
        for (long long g = g0; g <= g1; g++) {
            const long long f = first_unit(g);
            if (first_unit(g + 1) == f) continue;  // an empty range wrote nothing
            const float *pc = p.ws + (g * 2 + (f >= lo ? 0 : 1)) * (long long)TE + base;
#pragma unroll
            for (int j = 0; j < 8; j++) acc[j] += pc[j * 256];
        }
        float old[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            int m, n, nl;
            const bool ok = coords(base + j * 256, m, n, nl);
            old[j] = (p.beta != 0.0f && ok) ? p.out[out_addr<OP>(p, x, m, n)] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            int m, n, nl;
            if (!coords(base + j * 256, m, n, nl)) continue;
            float ev[5] = {0.f, 0.f, 1.f, 1.f, 0.f};
            if (FW) {
                if (staged) {
#pragma unroll
                    for (int k = 0; k < 5; k++) ev[k] = s_epi[k * BN + nl];
                } else {
                    if (p.bias) ev[0] = p.bias[n];
                    if (p.mean) {
                        ev[1] = p.mean[n];
                        ev[2] = p.inv_std[n];
                        ev[3] = p.gamma[n];
                        ev[4] = p.bnb[n];
                    }
                }
            }
            put_out<false>(p.out, out_addr<OP>(p, x, m, n), epi_value<OP>(p, acc[j], old[j], ev));
        }
    }
}

__global__ void scale_kernel(float *out, long long n, float beta) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = beta == 0.0f ? 0.0f : beta * out[i];
}

// One 64x64 tile of a 2-D transpose, dst[c*tc + r] = src[r*sr + c] for
// r0 <= r < min(r0 + 64, R), c0 <= c < min(c0 + 64, C), through LDS: 256
// threads, each with its 16 loads in flight at once (64 consecutive elements
// of a row per wave), then 16 stores along the transposed rows.  Every load
// is issued (an element outside the tensor reads element 0 and is dropped):
// a load guarded by a branch, with its LDS store right behind it, made the
// compiler wait for each load in turn.  The row pitch of 65 keeps the column
// reads on distinct banks.
constexpr int TT = 64;
template <typename E>
__device__ __attribute__((always_inline)) inline void transpose_tile(const E *__restrict__ src, E *__restrict__ dst,
                                                                     int R, int C, long long sr, long long tc, int r0,
                                                                     int c0, E (*t)[TT + 1]) {
    const int tx = (int)threadIdx.x & (TT - 1), ty = (int)threadIdx.x >> 6;
    const int c = c0 + tx;
    E v[TT / 4];
#pragma unroll
    for (int i = 0; i < TT / 4; i++) {
        const int r = r0 + 4 * i + ty;
        v[i] = src[(r < R && c < C) ? (long long)r * sr + c : 0];
    }
#pragma unroll
    for (int i = 0; i < TT / 4; i++) t[4 * i + ty][tx] = v[i];
    __syncthreads();
    const int r = r0 + tx;
#pragma unroll
    for (int i = 0; i < TT; i += 4) {
        const int cc = c0 + i + ty;
        if (r < R && cc < C) dst[(long long)cc * tc + r] = t[tx][i + ty];
    }
}

// dst[z*tz + c*tc + r] = src[z*sz + r*sr + c] for r < R, c < Cc, z < gridDim.z:
// a batch of 2-D transposes, one 64x64 tile per workgroup.
template <typename E>
__global__ void __launch_bounds__(256) strided_transpose_kernel(const E *__restrict__ src, E *__restrict__ dst, int R,
                                                                int Cc, long long sz, long long sr, long long tz,
                                                                long long tc) {
    __shared__ E t[TT][TT + 1];
    const long long z = blockIdx.z;
    transpose_tile(src + z * sz, dst + z * tz, R, Cc, sr, tc, (int)blockIdx.y * TT, (int)blockIdx.x * TT, t);
}

// [d0][d1][d2] -> [d0][d2][d1] for a short d2 <= PS_D2 (a kernel's taps),
// PS_ROWS rows of d1 at a time: the rows' PS_ROWS x d2 elements are one
// contiguous run of the input, loaded into LDS with coalesced loads, then
// stored as d2 runs of PS_ROWS consecutive elements of the output.  (A 64x64
// tile would leave most of its threads idle on d2 = 9.)  Item `it` of a
// permute with rb = ceil(d1 / PS_ROWS) row blocks per slice is slice it / rb.
constexpr int PS_ROWS = 256, PS_D2 = 16;
template <typename E>
__device__ __attribute__((always_inline)) inline void permute_short_item(const E *__restrict__ src,
                                                                         E *__restrict__ dst, int d1, int d2,
                                                                         long long it, E *s) {
    const int rb = (d1 + PS_ROWS - 1) / PS_ROWS;
    const long long a = it / rb;
    const int c0 = (int)(it - a * rb) * PS_ROWS;
    const int nc = min(PS_ROWS, d1 - c0);
    const long long slice = (long long)d1 * d2;
    const E *sp = src + a * slice + (long long)c0 * d2;
    E v[PS_D2];  // every load issued before the first LDS store (see transpose_tile)
#pragma unroll
    for (int k = 0; k < PS_D2; k++) {
        const int i = (int)threadIdx.x + 256 * k;
        v[k] = sp[i < nc * d2 ? i : 0];
    }
#pragma unroll
    for (int k = 0; k < PS_D2; k++) {
        const int i = (int)threadIdx.x + 256 * k;
        if (i < nc * d2) s[i] = v[k];
    }
    __syncthreads();
    E *dp = dst + a * slice + c0;
    const int c = (int)threadIdx.x;
    if (c < nc)
        for (int t = 0; t < d2; t++) dp[(long long)t * d1 + c] = s[c * d2 + t];
}

template <typename E>
__global__ void __launch_bounds__(256) permute_short_kernel(const E *__restrict__ src, E *__restrict__ dst, int d1,
                                                            int d2) {
    __shared__ E s[PS_ROWS * PS_D2];
    permute_short_item(src, dst, d1, d2, (long long)blockIdx.x, s);
}

// ---- a batch's element-wise jobs (cg_ewjob) in one launch.  A repack is a
// batch of Z 2-D transposes, src (z, r, c) at z*sz + r*sr + c to dst (z, c, r)
// at z*tz + c*tc + r, run as 32x32 tiles through LDS (cg_launch_permute3's
// decomposition); an item is one tile, or CG_EW_BLOCK elements of a scaling.
struct EwGeom {
    int R, C, Z;
    long long sz, sr, tz, tc;
};

__host__ __device__ inline EwGeom ew_geom(const cg_ewjob &j) {
    EwGeom g;
    const long long d0 = j.d0, d1 = j.d1, d2 = j.d2;
    if (j.kind == 0) {  // [d0][d1][d2] -> [d2][d1][d0]
        g.R = j.d0; g.C = j.d2; g.Z = j.d1;
        g.sz = d2; g.sr = d1 * d2; g.tz = d0; g.tc = d1 * d0;
    } else {            // [d0][d1][d2] -> [d0][d2][d1]
        g.R = j.d1; g.C = j.d2; g.Z = j.d0;
        g.sz = d1 * d2; g.sr = d2; g.tz = d2 * d1; g.tc = d1;
    }
    return g;
}

// kind 1 with a short d2 runs as permute_short_item
__host__ __device__ inline bool ew_short(const cg_ewjob &j) { return j.kind == 1 && j.d2 <= PS_D2; }

extern "C" long long cg_ew_items(const cg_ewjob *j) {
    if (j->kind >= 2) return (j->n + CG_EW_BLOCK - 1) / CG_EW_BLOCK;
    if (ew_short(*j)) return (long long)j->d0 * ((j->d1 + PS_ROWS - 1) / PS_ROWS);
    const EwGeom g = ew_geom(*j);
    return (long long)g.Z * ((g.R + TT - 1) / TT) * ((g.C + TT - 1) / TT);
}

template <typename E>
__global__ void __launch_bounds__(256) ew_grouped_kernel(const cg_ewjob *__restrict__ jobs,
                                                         const int *__restrict__ start, int njobs, int total) {
    __shared__ E t[TT][TT + 1];
    static_assert(PS_ROWS * PS_D2 <= TT * (TT + 1), "the short permute stages in the same LDS");
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = njobs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const cg_ewjob &j = jobs[lo];
        const int local = item - start[lo];
        if (j.kind == 2) {
            float *out = (float *)j.dst;
            const long long i0 = (long long)local * CG_EW_BLOCK;
            const long long i1 = min(i0 + CG_EW_BLOCK, j.n);
            for (long long i = i0 + threadIdx.x; i < i1; i += 256) out[i] = j.beta == 0.0f ? 0.0f : j.beta * out[i];
            continue;
        }
        if (j.kind == 3) {  // [d0][d1][d2] -> [d0][d1 - 1 - t][d2]
            const E *src = (const E *)j.src;
            E *dst = (E *)j.dst;
            const int plane = j.d1 * j.d2;
            const int i0 = local * CG_EW_BLOCK, i1 = (int)min((long long)i0 + CG_EW_BLOCK, j.n);
            for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
                const int a = i / plane, rem = i - a * plane;
                const int t = rem / j.d2, k = rem - t * j.d2;
                dst[i] = src[a * plane + (j.d1 - 1 - t) * j.d2 + k];
            }
            continue;
        }
        __syncthreads();  // the previous item's reads of the LDS are done
        if (ew_short(j)) {
            permute_short_item((const E *)j.src, (E *)j.dst, j.d1, j.d2, local, &t[0][0]);
            continue;
        }
        const EwGeom g = ew_geom(j);
        const int tr = (g.R + TT - 1) / TT, tcn = (g.C + TT - 1) / TT;
        const int z = local / (tr * tcn), rem = local - z * (tr * tcn);
        const int r0 = (rem / tcn) * TT, c0 = (rem - (rem / tcn) * tcn) * TT;
        transpose_tile((const E *)j.src + (long long)z * g.sz, (E *)j.dst + (long long)z * g.tz, g.R, g.C, g.sr,
                       g.tc, r0, c0, t);
    }
}

// [d0][d1][d2] -> [d0][d1 - 1 - t][d2] (a kernel's taps reversed): element i
// of the output reads its mirror row of the input.
template <typename E>
__global__ void reverse_mid_kernel(const E *__restrict__ src, E *__restrict__ dst, int d1, int d2, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long plane = (long long)d1 * d2;
    for (; i < n; i += stride) {
        const long long a = i / plane;
        const int rem = (int)(i - a * plane);
        const int t = rem / d2, k = rem - t * d2;
        dst[i] = src[a * plane + (long long)(d1 - 1 - t) * d2 + k];
    }
}

// ---- FP32 -> BF16 (round to nearest even) ----
__device__ inline uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x0040u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__global__ void to_bf16_kernel(const float *__restrict__ src, uint16_t *__restrict__ dst, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = f32_to_bf16_rne(src[i]);
}

// ---- per-tensor amax, then FP32 -> FP8 e4m3fn ----
__global__ void zero_u32_kernel(unsigned *p) { *p = 0u; }

__global__ void amax_kernel(const float *__restrict__ src, long long n, unsigned *amax_bits) {
    float m = 0.f;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) m = fmaxf(m, fabsf(src[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    __shared__ float wmax[16];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmaxf(m, wmax[w]);
        // non-negative floats order like their bit patterns
        atomicMax(amax_bits, __builtin_bit_cast(unsigned, m));
    }
}

// stored = e4m3fn(clamp(x / scale, -448, 448)), scale = amax / 448 (1 if amax == 0 or not finite)
__global__ void to_fp8_kernel(const float *__restrict__ src, uint8_t *__restrict__ dst, long long n,
                              const unsigned *amax_bits, float *scale_out) {
    const float amax = __builtin_bit_cast(float, *amax_bits);
    const float scale = (amax > 0.f && amax < 3.0e38f) ? amax / 448.0f : 1.0f;
    const float inv = 1.0f / scale;
    if (blockIdx.x == 0 && threadIdx.x == 0) *scale_out = scale;
    long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 2;
    const long long stride = (long long)gridDim.x * blockDim.x * 2;
    for (; i < n; i += stride) {
        const float a = fminf(fmaxf(src[i] * inv, -448.0f), 448.0f);
        const float b = (i + 1 < n) ? fminf(fmaxf(src[i + 1] * inv, -448.0f), 448.0f) : 0.0f;
        const int packed = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
        dst[i] = (uint8_t)(packed & 0xff);
        if (i + 1 < n) dst[i + 1] = (uint8_t)((packed >> 8) & 0xff);
    }
}

template <int OP>
static void launch_reduce(const cg_kparams *kp, dim3 grid, hipStream_t st) {
    hipLaunchKernelGGL(splitk_reduce_kernel<OP>, grid, dim3(256), 0, st, *kp);
}

}  // namespace cg

using namespace cg;

extern "C" int cg_igemm_sk_occupancy(int elem, int cfg, const cg_kparams *kp0) {
    cg_tilecfg t;
    if (cg_get_tilecfg(elem, cfg, &t)) return 0;
    if (kp0->op != CG_FWD_NHWC && kp0->op != CG_FWD_NCHW && kp0->op != CG_WGRAD_NHWC && kp0->op != CG_WGRAD_NCHW)
        return 0;
    cg_kparams k = *kp0;
    k.sk = -1;
    return cg_igemm_occupancy(elem, cfg, &k);
}

extern "C" int cg_launch_sk_fixup(const cg_kparams *kp, int BM, int BN, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    const int tiles_all = kp->mtiles * kp->ntiles;
    if (tiles_all <= kp->sk_dp) return 0;
    const unsigned tiles = (unsigned)(tiles_all - kp->sk_dp);  // the stream-K region
    int lbm = 0, lbn = 0;
    while ((1 << lbm) < BM) lbm++;
    while ((1 << lbn) < BN) lbn++;
    if ((1 << lbm) != BM || (1 << lbn) != BN || BM * BN < FIX_ELEMS) return (int)hipErrorInvalidValue;
    const unsigned blocks = tiles * (unsigned)(BM * BN / FIX_ELEMS);
#define CG_FX(OPK) hipLaunchKernelGGL(sk_fixup_kernel<OPK>, dim3(blocks), dim3(256), 0, st, *kp, lbm, lbn)
    switch (kp->op) {
    case CG_FWD_NHWC: CG_FX(CG_FWD_NHWC); break;
    case CG_FWD_NCHW: CG_FX(CG_FWD_NCHW); break;
    case CG_WGRAD_NHWC: CG_FX(CG_WGRAD_NHWC); break;
    case CG_WGRAD_NCHW: CG_FX(CG_WGRAD_NCHW); break;
    default: return (int)hipErrorInvalidValue;
    }
#undef CG_FX
    return (int)hipGetLastError();
}

extern "C" long long cg_reduce_items(const cg_kparams *kp) {
    const int per = 256 / reduce_groups(*kp);
    return (long long)kp->ncls_h * kp->ncls_w * (((long long)kp->M * kp->N + per - 1) / per);
}

extern "C" int cg_launch_splitk_reduce_grouped(int op, const cg_group *g, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    if (g->total <= 0) return 0;
#define CG_RG(OPK)                                                                                               \
    case OPK:                                                                                                    \
        hipLaunchKernelGGL(splitk_reduce_grouped_kernel<OPK>, dim3((unsigned)g->grid), dim3(256), 0, st, g->plist, \
                           g->start, g->nprob, g->total);                                                        \
        break;
    switch (op) {
        CG_RG(CG_FWD_NHWC)
        CG_RG(CG_FWD_NCHW)
        CG_RG(CG_WGRAD_NHWC)
        CG_RG(CG_WGRAD_NCHW)
        CG_RG(CG_DGRAD_NHWC)
        CG_RG(CG_DGRAD_NCHW)
    default: return (int)hipErrorInvalidValue;
    }
#undef CG_RG
    return (int)hipGetLastError();
}





extern "C" int cg_launch_splitk_reduce(const cg_kparams *kp, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    const long long total = (long long)kp->M * kp->N;  // rows of the largest class bound every class
    const int per = 256 / reduce_groups(*kp);  // outputs per block
    dim3 grid((unsigned)((total + per - 1) / per), kp->ncls_h * kp->ncls_w);
    switch (kp->op) {
    case CG_FWD_NHWC: launch_reduce<CG_FWD_NHWC>(kp, grid, st); break;
    case CG_FWD_NCHW: launch_reduce<CG_FWD_NCHW>(kp, grid, st); break;
    case CG_WGRAD_NHWC: launch_reduce<CG_WGRAD_NHWC>(kp, grid, st); break;
    case CG_WGRAD_NCHW: launch_reduce<CG_WGRAD_NCHW>(kp, grid, st); break;
    case CG_DGRAD_NHWC: launch_reduce<CG_DGRAD_NHWC>(kp, grid, st); break;
    case CG_DGRAD_NCHW: launch_reduce<CG_DGRAD_NCHW>(kp, grid, st); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

extern "C" int cg_launch_splitk_reduce_mixed(const cg_group *g, void *stream) {
    if (g->total <= 0) return 0;
    hipLaunchKernelGGL(splitk_reduce_mixed_kernel, dim3((unsigned)g->grid), dim3(256), 0, (hipStream_t)stream,
                       g->plist, g->start, g->nprob, g->total);
    return (int)hipGetLastError();
}

extern "C" int cg_launch_ew_grouped(const cg_ewjob *jobs, const int *start, int njobs, int total, int grid,
                                    int esize, void *stream) {
    if (total <= 0 || njobs <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    switch (esize) {
    case 4: hipLaunchKernelGGL(ew_grouped_kernel<float>, dim3((unsigned)grid), dim3(256), 0, st, jobs, start, njobs,
                               total); break;
    case 2: hipLaunchKernelGGL(ew_grouped_kernel<uint16_t>, dim3((unsigned)grid), dim3(256), 0, st, jobs, start,
                               njobs, total); break;
    case 1: hipLaunchKernelGGL(ew_grouped_kernel<uint8_t>, dim3((unsigned)grid), dim3(256), 0, st, jobs, start,
                               njobs, total); break;
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

template <typename E>
static int launch_transpose(const void *src, void *dst, int R, int Cc, int batch, long long sz, long long sr,
                            long long tz, long long tc, hipStream_t st) {
    dim3 grid((Cc + TT - 1) / TT, (R + TT - 1) / TT, batch);
    hipLaunchKernelGGL(strided_transpose_kernel<E>, grid, dim3(256), 0, st, (const E *)src, (E *)dst, R, Cc, sz, sr,
                       tz, tc);
    return (int)hipGetLastError();
}

extern "C" int cg_launch_permute3(const void *src, void *dst, int d0, int d1, int d2, int perm, int esize,
                                  void *stream) {
    if (d0 <= 0 || d1 <= 0 || d2 <= 0) return 0;
    if (perm == 2) {  // [d0][d1][d2] -> [d0][d1 reversed][d2]: rows of d2 elements move whole
        const long long n = (long long)d0 * d1 * d2;
        long long blocks = (n + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipStream_t st = (hipStream_t)stream;
        switch (esize) {
        case 4: hipLaunchKernelGGL(reverse_mid_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, st,
                                   (const float *)src, (float *)dst, d1, d2, n); break;
        case 2: hipLaunchKernelGGL(reverse_mid_kernel<uint16_t>, dim3((unsigned)blocks), dim3(256), 0, st,
                                   (const uint16_t *)src, (uint16_t *)dst, d1, d2, n); break;
        case 1: hipLaunchKernelGGL(reverse_mid_kernel<uint8_t>, dim3((unsigned)blocks), dim3(256), 0, st,
                                   (const uint8_t *)src, (uint8_t *)dst, d1, d2, n); break;
        default: return (int)hipErrorInvalidValue;
        }
        return (int)hipGetLastError();
    }
    int R, Cc, batch;
    long long sz, sr, tz, tc;
    if (perm == 0) {  // [d0][d1][d2] -> [d2][d1][d0]: batch over d1
        R = d0; Cc = d2; batch = d1;
        sz = d2; sr = (long long)d1 * d2; tz = d0; tc = (long long)d1 * d0;
    } else {          // [d0][d1][d2] -> [d0][d2][d1]: batch over d0
        if (d2 <= PS_D2) {
            const long long blocks = (long long)d0 * ((d1 + PS_ROWS - 1) / PS_ROWS);
            if (blocks > 0x7fffffffLL) return (int)hipErrorInvalidValue;
            hipStream_t st = (hipStream_t)stream;
            switch (esize) {
            case 4: hipLaunchKernelGGL(permute_short_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, st,
                                       (const float *)src, (float *)dst, d1, d2); break;
            case 2: hipLaunchKernelGGL(permute_short_kernel<uint16_t>, dim3((unsigned)blocks), dim3(256), 0, st,
                                       (const uint16_t *)src, (uint16_t *)dst, d1, d2); break;
            case 1: hipLaunchKernelGGL(permute_short_kernel<uint8_t>, dim3((unsigned)blocks), dim3(256), 0, st,
                                       (const uint8_t *)src, (uint8_t *)dst, d1, d2); break;
            default: return (int)hipErrorInvalidValue;
            }
            return (int)hipGetLastError();
        }
        R = d1; Cc = d2; batch = d0;
        sz = (long long)d1 * d2; sr = d2; tz = (long long)d2 * d1; tc = d1;
    }
    if (batch > 65535) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    switch (esize) {
    case 4: return launch_transpose<float>(src, dst, R, Cc, batch, sz, sr, tz, tc, st);
    case 2: return launch_transpose<uint16_t>(src, dst, R, Cc, batch, sz, sr, tz, tc, st);
    case 1: return launch_transpose<uint8_t>(src, dst, R, Cc, batch, sz, sr, tz, tc, st);
    }
    return (int)hipErrorInvalidValue;
}

extern "C" int cg_launch_transpose2d(const void *src, long long lds, void *dst, long long ldd, int rows, int cols,
                                     int esize, void *stream) {
    if (rows <= 0 || cols <= 0) return 0;
    if ((rows + TT - 1) / TT > 65535) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    switch (esize) {
    case 4: return launch_transpose<float>(src, dst, rows, cols, 1, 0, lds, 0, ldd, st);
    case 2: return launch_transpose<uint16_t>(src, dst, rows, cols, 1, 0, lds, 0, ldd, st);
    case 1: return launch_transpose<uint8_t>(src, dst, rows, cols, 1, 0, lds, 0, ldd, st);
    }
    return (int)hipErrorInvalidValue;
}

// FP32 -> BF16 (scale = 1) or FP8 e4m3fn with a per-tensor amax scale.
// `scratch` holds 8 device bytes (amax bits + scale); `scale_dev` receives the scale.
extern "C" int cg_launch_quantize(const float *src, void *dst, long long n, int elem, void *scratch, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    if (n <= 0) return 0;
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (elem == CG_ELEM_BF16) {
        hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, (uint16_t *)dst, n);
        return (int)hipGetLastError();
    }
    if (elem != CG_ELEM_FP8) return (int)hipErrorInvalidValue;
    unsigned *amax = (unsigned *)scratch;
    float *scale = (float *)scratch + 1;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, st, amax);
    hipLaunchKernelGGL(amax_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, n, amax);
    hipLaunchKernelGGL(to_fp8_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, (uint8_t *)dst, n, amax, scale);
    return (int)hipGetLastError();
}
