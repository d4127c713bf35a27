// transforms.hip -- explicit convolution transforms and plain copies.
//
// The reference materialises its patch matrices only in its tests and in the
// explicit path (src/im2row_nhwc.c:147-229, src/im2col_nchw.c:146-233,
// :206-233, :337-359); here they are bandwidth-bound grid-stride kernels:
//   im2row / im2col: 4- or 2-byte elements, threads along the patch-matrix
//     row so the stores coalesce; only in-bounds taps are written (the caller
//     zero-fills, as the reference's tests do);
//   row2im / col2im: a gather per image element over the taps that reach it
//     (no atomics, fixed order), added to the image;
//   transpose_nchw: [kn-block][b][ho][wo] -> [b][kn][ho][wo] with beta;
//   copy2d: strided 2-D copy for packing GEMM operands.
#include <hip/hip_runtime.h>

#include "../cg_internal.h"

namespace {

constexpr int TB = 256;

inline unsigned grid_for(long long n) {
    long long g = (n + TB - 1) / TB;
    if (g > 65536) g = 65536;
    return (unsigned)(g > 0 ? g : 1);
}

// rows (NHWC: P x K, ld) from the image x: a block walks patch rows p
// (grid-stride), its threads the K = c*kh*kw elements of a row (r = (c*kh +
// kx)*kw + ky), so the stores of a row coalesce; the pixel decomposition is
// done once per row and the per-element index math is 32-bit.  E is the
// element type (4 or 2 bytes: the copy does no arithmetic).
template <typename E>
__global__ void im2row_nhwc_kernel(cg_tshape s, const E *__restrict__ x, E *__restrict__ out, long long ld) {
    const int taps = s.kh * s.kw, K = s.c * taps, hw = s.ho * s.wo;
    const long long P = (long long)s.b * hw;
    for (long long p = blockIdx.x; p < P; p += gridDim.x) {
        const int b = (int)(p / hw), rem = (int)(p - (long long)b * hw);
        const int xo = rem / s.wo, yo = rem - xo * s.wo;
        const int ix0 = s.vs * xo - s.vp, iy0 = s.hs * yo - s.hp;
        const E *img = x + (long long)b * s.h * s.w * s.c;
        E *row = out + p * ld;
        for (int r = threadIdx.x; r < K; r += blockDim.x) {
            const int c = r / taps, t = r - c * taps;
            const int kx = t / s.kw, ky = t - kx * s.kw;
            const int ix = ix0 + s.vd * kx, iy = iy0 + s.hd * ky;
            if ((unsigned)ix < (unsigned)s.h && (unsigned)iy < (unsigned)s.w) row[r] = img[((long long)ix * s.w + iy) * s.c + c];
        }
    }
}

// cols (NCHW: K x P, ld) from the image x: a block walks the pixels of one
// patch row r (grid-stride over (r, pixel block)), threads along p, so the
// stores coalesce and the loads run along an image row.
template <typename E>
__global__ void im2col_nchw_kernel(cg_tshape s, const E *__restrict__ x, E *__restrict__ out, long long ld) {
    const int taps = s.kh * s.kw, K = s.c * taps, hw = s.ho * s.wo;
    const long long P = (long long)s.b * hw;
    const long long pblocks = (P + blockDim.x - 1) / blockDim.x;
    for (long long w = blockIdx.x; w < (long long)K * pblocks; w += gridDim.x) {
        const int r = (int)(w / pblocks);
        const long long p = (w - (long long)r * pblocks) * blockDim.x + threadIdx.x;
        if (p >= P) continue;
        const int c = r / taps, t = r - c * taps;
        const int kx = t / s.kw, ky = t - kx * s.kw;
        const int b = (int)(p / hw), rem = (int)(p - (long long)b * hw);
        const int xo = rem / s.wo, yo = rem - xo * s.wo;
        const int ix = s.vs * xo + s.vd * kx - s.vp, iy = s.hs * yo + s.hd * ky - s.hp;
        if ((unsigned)ix < (unsigned)s.h && (unsigned)iy < (unsigned)s.w)
            out[(long long)r * ld + p] = x[(((long long)b * s.c + c) * s.h + ix) * s.w + iy];
    }
}

// x += row2im(rows) / col2im(cols): image element e gathers every (pixel, tap)
// whose input coordinate is e.
__global__ void row2im_kernel(cg_tshape s, const float *__restrict__ in, long long ld, float *__restrict__ x,
                              int nchw) {
    const long long total = (long long)s.b * s.h * s.w * s.c;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        int b, c, ix, iy;
        if (nchw) {
            iy = (int)(e % s.w);
            long long q = e / s.w;
            ix = (int)(q % s.h);
            q /= s.h;
            c = (int)(q % s.c);
            b = (int)(q / s.c);
        } else {
            c = (int)(e % s.c);
            long long q = e / s.c;
            iy = (int)(q % s.w);
            q /= s.w;
            ix = (int)(q % s.h);
            b = (int)(q / s.h);
        }
        float sum = 0.f;
        for (int kx = 0; kx < s.kh; kx++) {
            const int tx = ix + s.vp - s.vd * kx;
            if (tx < 0 || tx % s.vs) continue;
            const int xo = tx / s.vs;
            if (xo >= s.ho) continue;
            for (int ky = 0; ky < s.kw; ky++) {
                const int ty = iy + s.hp - s.hd * ky;
                if (ty < 0 || ty % s.hs) continue;
                const int yo = ty / s.hs;
                if (yo >= s.wo) continue;
                const long long p = ((long long)b * s.ho + xo) * s.wo + yo;
                const long long r = ((long long)c * s.kh + kx) * s.kw + ky;
                sum += nchw ? in[r * ld + p] : in[p * ld + r];
            }
        }
        x[e] += sum;
    }
}

// out[((b*kn + start_col + j)*ho + x)*wo + y] (=|+= beta*out) in[j*ld + i],
// (b, x, y) the pixel start_row + i.
__global__ void transpose_nchw_kernel(int rows, int cols, const float *__restrict__ in, long long ld, float beta,
                                      float *__restrict__ out, int kn, int ho, int wo, int start_row, int start_col) {
    const long long total = (long long)rows * cols;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e / rows), i = (int)(e - (long long)j * rows);
        const long long p = (long long)start_row + i;
        const long long hw = (long long)ho * wo;
        const long long b = p / hw, rem = p - b * hw;
        const long long idx = (b * kn + start_col + j) * hw + rem;
        const float v = in[(long long)j * ld + i];
        out[idx] = beta == 0.0f ? v : beta * out[idx] + v;
    }
}

template <typename E>
__global__ void copy2d_kernel(const E *__restrict__ src, long long lds, E *__restrict__ dst, long long ldd,
                              long long rows, long long cols) {
    const long long total = rows * cols;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long r = e / cols, c = e - r * cols;
        dst[r * ldd + c] = src[r * lds + c];
    }
}

}  // namespace

extern "C" int cg_launch_im2row(const cg_tshape *s, const void *x, void *out, long long ld, int nchw, int esize,
                                void *stream) {
    const long long P = (long long)s->b * s->ho * s->wo, K = (long long)s->c * s->kh * s->kw;
    if (P <= 0 || K <= 0) return 0;
    if (esize != 4 && esize != 2) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    if (!nchw) {
        const unsigned g = (unsigned)(P < 65536 ? P : 65536);
        if (esize == 4)
            hipLaunchKernelGGL(im2row_nhwc_kernel<uint32_t>, dim3(g), dim3(TB), 0, st, *s, (const uint32_t *)x,
                               (uint32_t *)out, ld);
        else
            hipLaunchKernelGGL(im2row_nhwc_kernel<uint16_t>, dim3(g), dim3(TB), 0, st, *s, (const uint16_t *)x,
                               (uint16_t *)out, ld);
    } else {
        const unsigned g = grid_for(P * K);
        if (esize == 4)
            hipLaunchKernelGGL(im2col_nchw_kernel<uint32_t>, dim3(g), dim3(TB), 0, st, *s, (const uint32_t *)x,
                               (uint32_t *)out, ld);
        else
            hipLaunchKernelGGL(im2col_nchw_kernel<uint16_t>, dim3(g), dim3(TB), 0, st, *s, (const uint16_t *)x,
                               (uint16_t *)out, ld);
    }
    return (int)hipGetLastError();
}

extern "C" int cg_launch_row2im(const cg_tshape *s, const float *in, long long ld, float *x, int nchw,
                                void *stream) {
    const long long n = (long long)s->b * s->h * s->w * s->c;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(row2im_kernel, dim3(grid_for(n)), dim3(TB), 0, (hipStream_t)stream, *s, in, ld, x, nchw);
    return (int)hipGetLastError();
}

extern "C" int cg_launch_transpose_nchw(int rows, int cols, const float *in, long long ld, float beta, float *out,
                                        int kn, int ho, int wo, int start_row, int start_col, void *stream) {
    const long long n = (long long)rows * cols;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(transpose_nchw_kernel, dim3(grid_for(n)), dim3(TB), 0, (hipStream_t)stream, rows, cols, in,
                       ld, beta, out, kn, ho, wo, start_row, start_col);
    return (int)hipGetLastError();
}

extern "C" int cg_launch_copy2d(const void *src, long long lds, void *dst, long long ldd, long long rows,
                                long long cols, int esize, void *stream) {
    const long long n = rows * cols;
    if (n <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    switch (esize) {
    case 4:
        hipLaunchKernelGGL(copy2d_kernel<uint32_t>, dim3(grid_for(n)), dim3(TB), 0, st, (const uint32_t *)src, lds,
                           (uint32_t *)dst, ldd, rows, cols);
        break;
    case 2:
        hipLaunchKernelGGL(copy2d_kernel<uint16_t>, dim3(grid_for(n)), dim3(TB), 0, st, (const uint16_t *)src, lds,
                           (uint16_t *)dst, ldd, rows, cols);
        break;
    case 1:
        hipLaunchKernelGGL(copy2d_kernel<uint8_t>, dim3(grid_for(n)), dim3(TB), 0, st, (const uint8_t *)src, lds,
                           (uint8_t *)dst, ldd, rows, cols);
        break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}
