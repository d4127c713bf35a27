// transforms.hip -- explicit convolution transforms and plain copies.
//
// The reference materialises its patch matrices only in its tests and in the
// explicit path (src/im2row_nhwc.c:147-229, src/im2col_nchw.c:146-233,
// :206-233, :337-359); here they are bandwidth-bound grid-stride kernels:
//   im2row / im2col: one thread per patch-matrix element, the element index
//     fastest along the matrix row so the stores coalesce; only in-bounds
//     taps are written (the caller zero-fills, as the reference's tests do);
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

// rows (NHWC: P x K, ld) or cols (NCHW: K x P, ld) from the image x.
__global__ void im2row_kernel(cg_tshape s, const float *__restrict__ x, float *__restrict__ out, long long ld,
                              int nchw) {
    const long long taps = (long long)s.kh * s.kw;
    const long long K = (long long)s.c * taps;
    const long long P = (long long)s.b * s.ho * s.wo;
    const long long total = P * K;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        long long p, r;
        if (nchw) { r = i / P; p = i - r * P; }
        else { p = i / K; r = i - p * K; }
        const int c = (int)(r / taps), t = (int)(r - (long long)c * taps);
        const int kx = t / s.kw, ky = t - kx * s.kw;
        const int hw = s.ho * s.wo;
        const int b = (int)(p / hw), rem = (int)(p - (long long)b * hw);
        const int xo = rem / s.wo, yo = rem - xo * s.wo;
        const int ix = s.vs * xo + s.vd * kx - s.vp, iy = s.hs * yo + s.hd * ky - s.hp;
        if (ix < 0 || ix >= s.h || iy < 0 || iy >= s.w) continue;
        if (nchw) out[r * ld + p] = x[(((long long)b * s.c + c) * s.h + ix) * s.w + iy];
        else out[p * ld + r] = x[(((long long)b * s.h + ix) * s.w + iy) * s.c + c];
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

extern "C" int cg_launch_im2row(const cg_tshape *s, const float *x, float *out, long long ld, int nchw,
                                void *stream) {
    const long long n = (long long)s->b * s->ho * s->wo * s->c * s->kh * s->kw;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(im2row_kernel, dim3(grid_for(n)), dim3(TB), 0, (hipStream_t)stream, *s, x, out, ld, nchw);
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
