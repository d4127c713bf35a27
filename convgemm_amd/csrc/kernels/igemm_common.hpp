// igemm_common.hpp -- device helpers shared by the implicit-GEMM kernels and
// the split-K reduction: fast integer division, the per-block problem
// context, and the output store with the fused epilogue.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../cg_internal.h"

namespace cg {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned char u8x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

// Operand element types: float (FP32), uint16_t (bfloat16 bits), uint8_t
// (OCP fp8 e4m3fn bits).  VEC elements make one 16-byte chunk; G is the
// reduction length one MFMA group covers per lane pair (f32: 8 x 32x32x2,
// bf16: 32x32x16, fp8: 32x32x64).
template <typename T> struct ET;
template <> struct ET<float> {
    static constexpr int VEC = 4, G = 16;
    typedef f32x4 V;
};
template <> struct ET<uint16_t> {
    static constexpr int VEC = 8, G = 16;
    typedef u16x8 V;
};
template <> struct ET<uint8_t> {
    static constexpr int VEC = 16, G = 64;
    typedef u8x16 V;
};

// Buffer-resource loads: an offset at or beyond the resource size returns 0
// in hardware, so masked gathers need neither branches nor selects.
// Operand tensors are limited to < 2 GiB per launch (the host splits larger
// problems along the batch); OOB is an offset no resource reaches.
constexpr unsigned OOB = 0x80000000u;

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void *p, long long bytes) {
    unsigned n = bytes >= (long long)OOB ? OOB : (unsigned)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)n, 0x00020000);
}

template <typename T>
__device__ inline typename ET<T>::V ld_vec(__amdgpu_buffer_rsrc_t r, unsigned boff) {
    return __builtin_bit_cast(typename ET<T>::V, __builtin_amdgcn_raw_buffer_load_b128(r, boff, 0, 0));
}
template <typename T> __device__ inline T ld_one(__amdgpu_buffer_rsrc_t r, unsigned boff);
template <> __device__ inline float ld_one<float>(__amdgpu_buffer_rsrc_t r, unsigned boff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, boff, 0, 0));
}
template <> __device__ inline uint16_t ld_one<uint16_t>(__amdgpu_buffer_rsrc_t r, unsigned boff) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, boff, 0, 0);
}
template <> __device__ inline uint8_t ld_one<uint8_t>(__amdgpu_buffer_rsrc_t r, unsigned boff) {
    return __builtin_amdgcn_raw_buffer_load_b8(r, boff, 0, 0);
}

// Division by a run-time constant d >= 1 for 0 <= n < 2^31
// (Granlund-Montgomery round-up method): q = (n * m) >> s with
// s = 31 + ceil(log2 d), m = ceil(2^s / d) < 2^32.
struct FastDiv {
    uint32_t m;
    uint32_t s;
    int d;
};

__device__ __host__ inline FastDiv make_fastdiv(int d) {
    uint32_t l = 0;
    while ((1u << l) < (uint32_t)d) l++;
    uint32_t s = 31 + l;
    uint64_t m = ((1ull << s) + (uint64_t)d - 1) / (uint64_t)d;
    FastDiv f;
    f.m = (uint32_t)m;
    f.s = s;
    f.d = d;
    return f;
}

__device__ __host__ inline int fdiv(int n, const FastDiv &f) {
    return (int)(((uint64_t)(uint32_t)n * f.m) >> f.s);
}

// Problem facts a block needs, computed once in the prologue.
struct Ctx {
    FastDiv howo, wo, cls_hw, cls_w, c;
    int rh, rw;      // DGRAD residue class
    int Hc, Wc;      // DGRAD class extents
    int Mrows;       // valid logical rows of this block's problem/class
};

// The host's multiply-and-shift constants (cg_kparams.fdm_*, fds_*).
__device__ inline FastDiv host_fastdiv(unsigned m, int s, int d) {
    FastDiv f;
    f.m = m;
    f.s = (uint32_t)s;
    f.d = d;
    return f;
}

__device__ inline Ctx make_ctx(const cg_kparams &p, int cls) {
    Ctx x;
    int op = p.op;
    x.howo = host_fastdiv(p.fdm_howo, p.fds_howo, p.Ho * p.Wo);
    x.wo = host_fastdiv(p.fdm_wo, p.fds_wo, p.Wo);
    x.c = host_fastdiv(p.fdm_c, p.fds_c, p.C);
    x.rh = 0;
    x.rw = 0;
    x.Hc = p.Ho;
    x.Wc = p.Wo;
    if (op == CG_DGRAD_NHWC || op == CG_DGRAD_NCHW) {
        x.rh = cls / p.ncls_w;
        x.rw = cls - x.rh * p.ncls_w;
        x.Hc = (p.H - x.rh + p.vs - 1) / p.vs;
        x.Wc = (p.W - x.rw + p.hs - 1) / p.hs;
        x.Mrows = p.B * x.Hc * x.Wc;
    } else if (op == CG_WGRAD_NHWC || op == CG_WGRAD_NCHW) {
        x.Mrows = p.taps * p.C;
    } else {
        x.Mrows = (int)p.P;
    }
    if (op == CG_DGRAD_NHWC || op == CG_DGRAD_NCHW) {  // class extents differ from Ho, Wo
        x.cls_hw = make_fastdiv(x.Hc * x.Wc > 0 ? x.Hc * x.Wc : 1);
        x.cls_w = make_fastdiv(x.Wc > 0 ? x.Wc : 1);
    } else {
        x.cls_hw = x.howo;
        x.cls_w = x.wo;
    }
    return x;
}

// Write one finished accumulator to the output tensor, applying
// alpha/beta and (forward only) bias + BN + ReLU.
//   m: logical GEMM row (FWD: pixel; DGRAD: class pixel; WGRAD: tap*C + c)
//   n: logical GEMM col (FWD/WGRAD: kn; DGRAD: c)
// Epilogue order follows src/im2row_nhwc.c:276-282 and
// src/im2col_nchw.c:422-431.
// Element offset of output (m, n) in the operation's layout.
template <int OP>
__device__ inline long long out_addr(const cg_kparams &p, const Ctx &x, int m, int n) {
    if constexpr (OP == CG_FWD_NHWC) {
        return (long long)m * p.KN + n;
    } else if constexpr (OP == CG_FWD_NCHW) {
        int b = fdiv(m, x.howo);
        int pix = m - b * (p.Ho * p.Wo);
        return ((long long)b * p.KN + n) * (p.Ho * p.Wo) + pix;
    } else if constexpr (OP == CG_DGRAD_NHWC || OP == CG_DGRAD_NCHW) {
        int b = fdiv(m, x.cls_hw);
        int rem = m - b * (x.Hc * x.Wc);
        int i = fdiv(rem, x.cls_w);
        int j = rem - i * x.Wc;
        int ih = x.rh + p.vs * i, iw = x.rw + p.hs * j;
        if constexpr (OP == CG_DGRAD_NHWC)
            return (((long long)b * p.H + ih) * p.W + iw) * p.C + n;
        else
            return (((long long)b * p.C + n) * p.H + ih) * p.W + iw;
    } else {  // WGRAD
        int tap = fdiv(m, x.c);
        int c = m - tap * p.C;
        if constexpr (OP == CG_WGRAD_NHWC)
            return ((long long)c * p.taps + tap) * p.KN + n;
        else
            return (long long)n * ((long long)p.C * p.taps) + (long long)c * p.taps + tap;
    }
}

// The reference's epilogue order (src/im2row_nhwc.c:277-281): t = alpha*acc,
// t += beta*out, t += bias, t = (t - mean)*inv_std, t = t*gamma + bn_beta,
// ReLU.  `ev` holds the channel's (bias, mean, inv_std, gamma, bn_beta).
template <int OP>
__device__ inline float epi_value(const cg_kparams &p, float acc, float old, const float (&ev)[5]) {
    // Every product and sum rounded on its own: whether the compiler fuses a
    // multiply and an add depends on the code around the inlined call, and
    // every kernel that writes a given output must round it the same way.
#pragma clang fp contract(off)
    float t = p.alpha * p.scale * acc;
    if (p.beta != 0.0f) t += p.beta * old;
    if constexpr (OP == CG_FWD_NHWC || OP == CG_FWD_NCHW) {
        if (p.bias) t += ev[0];
        if (p.mean) {
            t = (t - ev[1]) * ev[2];
            t = t * ev[3] + ev[4];
        }
        if (p.relu && t < 0.0f) t = 0.0f;
    }
    return t;
}

// Store one final output element; NT: as a non-temporal store, which does
// not keep the line in L2.  The kernels never read their outputs again, and
// with BF16 and FP8 operands (the FP32 output is 2-4x the bytes of the
// input) the forward kernels ran 6-12% faster with the input rows that later
// taps gather again left in L2; FP32 forward tiles were 1-2% slower and the
// backward operations unchanged, so they store normally.
template <bool NT>
__device__ __attribute__((always_inline)) inline void put_out(float *out, long long addr, float v) {
    if constexpr (NT) __builtin_nontemporal_store(v, out + addr);
    else out[addr] = v;
}

template <int OP, bool NT = false>
__device__ inline void store_out(const cg_kparams &p, const Ctx &x, int m, int n, float acc) {
    const long long addr = out_addr<OP>(p, x, m, n);
    float ev[5] = {0.f, 0.f, 1.f, 1.f, 0.f};
    if constexpr (OP == CG_FWD_NHWC || OP == CG_FWD_NCHW) {
        if (p.bias) ev[0] = p.bias[n];
        if (p.mean) {
            ev[1] = p.mean[n];
            ev[2] = p.inv_std[n];
            ev[3] = p.gamma[n];
            ev[4] = p.bnb[n];
        }
    }
    put_out<NT>(p.out, addr, epi_value<OP>(p, acc, p.beta != 0.0f ? p.out[addr] : 0.f, ev));
}

// ---- tile write-back ----
// A load issued after a store waits for that store as well (vmcnt counts
// both), and the compiler cannot move loads of the epilogue vectors or of
// beta*out across stores to `out` (they may alias).  Per-element store_out in
// a tile's write-back therefore costs one memory round trip per output.  The
// tile write-back instead stages the forward epilogue's per-channel vectors
// of its columns in LDS once, and loads the beta term of 16 outputs before
// storing any of them.

// Does the tile write-back stage epilogue vectors (uniform per workgroup)?
template <int OP>
__device__ inline bool epi_staged(const cg_kparams &p) {
    return (OP == CG_FWD_NHWC || OP == CG_FWD_NCHW) && (p.bias || p.mean);
}

// Stage the vectors of channels [n0, n0 + cnt) as 5 rows of cnt floats at s
// (bias, mean, inv_std, gamma, bn_beta).  Called by all threads of the
// workgroup between barriers.  The five loads of a channel are all issued
// before its LDS stores: with each store right behind its own load, the
// compiler waited for the loads one group at a time.
__device__ inline void epi_stage(const cg_kparams &p, int n0, int cnt, float *s, int tid, int nthreads) {
    for (int i = tid; i < cnt; i += nthreads) {
        const int n = n0 + i;
        float v[5] = {0.f, 0.f, 1.f, 1.f, 0.f};
        if (n < p.KN) {
            if (p.bias) v[0] = p.bias[n];
            if (p.mean) {
                v[1] = p.mean[n];
                v[2] = p.inv_std[n];
                v[3] = p.gamma[n];
                v[4] = p.bnb[n];
            }
        }
#pragma unroll
        for (int k = 0; k < 5; k++) s[k * cnt + i] = v[k];
    }
}

// Write back one 32x32 MFMA block (16 accumulators per lane) with the fused
// epilogue.  pos(v, m, n, nl) gives accumulator v's output (m, n), its column
// within the staged vectors (nl), and whether the output exists.
template <int OP, bool NT = false, typename Pos>
__device__ __attribute__((always_inline)) inline void store_block(const cg_kparams &p, const Ctx &x, const f32x16 &acc,
                                                                   const float *s_epi, int cnt, Pos pos) {
    auto vec = [&](int nl, float (&ev)[5]) __attribute__((always_inline)) {
        ev[0] = 0.f, ev[1] = 0.f, ev[2] = 1.f, ev[3] = 1.f, ev[4] = 0.f;
        if (s_epi) {
#pragma unroll
            for (int k = 0; k < 5; k++) ev[k] = s_epi[k * cnt + nl];
        }
    };
    if (p.beta == 0.0f) {  // no loads at all: nothing for a store to wait behind
#pragma unroll
        for (int v = 0; v < 16; v++) {
            int m, n, nl;
            if (!pos(v, m, n, nl)) continue;
            float ev[5];
            vec(nl, ev);
            put_out<NT>(p.out, out_addr<OP>(p, x, m, n), epi_value<OP>(p, acc[v], 0.f, ev));
        }
        return;
    }
    // beta != 0: in groups of 4 outputs, the old values first, without branches
    // (an output that does not exist reads out[0] and is not stored), then the
    // stores -- one wait per group instead of one per output, in few registers
#pragma unroll
    for (int g = 0; g < 16; g += 4) {
        float old[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int m, n, nl;
            const bool ok = pos(g + u, m, n, nl);
            old[u] = p.out[ok ? out_addr<OP>(p, x, m, n) : 0];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int m, n, nl;
            if (!pos(g + u, m, n, nl)) continue;
            float ev[5];
            vec(nl, ev);
            put_out<NT>(p.out, out_addr<OP>(p, x, m, n), epi_value<OP>(p, acc[g + u], old[u], ev));
        }
    }
}

// Tile t -> (M panel mt, N panel nt) in the problem's raster order:
//   0: the N panels of one M panel adjacent (t = mt * ntiles + nt): the
//      gathered operand's panel is fetched once and the small weight panel
//      stays L2-resident;
//   1: the M panels of one N panel adjacent (t = nt * mtiles + mt): the
//      weight panel is fetched once per run of M panels (the reference's
//      A3B2C0 order, src/gemm_blis_A3B2C0.c:55-80, puts the B panel outside
//      its ic loop likewise);
//   2: bands of 8 M panels, the M panels of one N panel adjacent within a
//      band: an XCD's run of tiles touches 8 M panels and their N panels.
__device__ inline void tile_mn(const cg_kparams &p, int t, int &mt, int &nt) {
    if (p.raster == 1) {
        nt = t / p.mtiles;
        mt = t - nt * p.mtiles;
    } else if (p.raster == 2) {
        const int band = t / (8 * p.ntiles);
        const int m0 = band * 8;
        const int bh = min(8, p.mtiles - m0);
        const int r = t - band * 8 * p.ntiles;
        nt = r / bh;
        mt = m0 + (r - nt * bh);
    } else {
        mt = t / p.ntiles;
        nt = t - mt * p.ntiles;
    }
}

// ---- stream-K schedule (blocks of a *_sk_kernel) ----
// Units are (tile, K step) pairs numbered tile * KT + k.  Block b < sk_dp
// owns the units of data-parallel tile xcd_run(b, sk_dp); block sk_dp + i
// owns units [sk_dp*KT + g*U/G, sk_dp*KT + (g+1)*U/G), g = xcd_run(i, G),
// U = (tiles - sk_dp)*KT, G = p.sk.
//
// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8);
// xcd_run gives each XCD a contiguous run of the index space, so neighbouring
// tiles (shared weight panel, shared input halo) meet in one L2.
__device__ inline int xcd_run(int bid, int nb) {
    const int xcd = bid & 7, per = nb >> 3, rem = nb & 7;
    return xcd * per + min(xcd, rem) + (bid >> 3);
}

// The kernel's cg_kparams argument read through the kernarg segment pointer
// (loads from constant memory where they are used) instead of as values
// loaded at the kernel entry and kept live across the whole body.
// The argument must be the kernel's first (only) parameter, which sits at
// the start of the kernarg segment.  The host pass of the compiler only
// parses this device function.
__device__ inline const cg_kparams &kernarg_params(const cg_kparams &arg) {
#if defined(__HIP_DEVICE_COMPILE__)
    (void)arg;
    return *reinterpret_cast<const cg_kparams *>(__builtin_amdgcn_kernarg_segment_ptr());
#else
    return arg;
#endif
}

struct SkRange {
    long long u, u_first, u_end;
    int g;  // stream-K workgroup index (-1: a data-parallel block)
};

struct SkPiece {
    int tile, k0, k1;
    float *piece;  // nullptr: the whole K range, written back directly
};

__device__ inline SkRange sk_range(const cg_kparams &p) {
    SkRange r;
    const int b = (int)blockIdx.x;
    if (b < p.sk_dp) {
        r.g = -1;
        r.u_first = (long long)xcd_run(b, p.sk_dp) * p.KT;
        r.u_end = r.u_first + p.KT;
    } else {
        r.g = xcd_run(b - p.sk_dp, p.sk);
        const long long base = (long long)p.sk_dp * p.KT;
        const long long U = ((long long)p.mtiles * p.ntiles - p.sk_dp) * p.KT;
        r.u_first = base + (long long)r.g * U / p.sk;
        r.u_end = base + (long long)(r.g + 1) * U / p.sk;
    }
    r.u = r.u_first;
    return r;
}

// The next piece of the range (false when it is exhausted).  A piece that is
// not a whole tile goes to the workgroup's slot 0 (its first piece) or 1.
__device__ inline bool sk_next(const cg_kparams &p, SkRange &r, SkPiece &q, int tile_elems) {
    if (r.u >= r.u_end) return false;
    const int t = (int)(r.u / p.KT);
    const int k0 = (int)(r.u - (long long)t * p.KT);
    const int k1 = (int)min((long long)p.KT, (long long)k0 + (r.u_end - r.u));
    q.tile = t;
    q.k0 = k0;
    q.k1 = k1;
    q.piece = (k0 == 0 && k1 == p.KT) ? nullptr
                                      : p.ws + ((long long)r.g * 2 + (r.u == r.u_first ? 0 : 1)) * tile_elems;
    r.u += k1 - k0;
    return true;
}

}  // namespace cg
