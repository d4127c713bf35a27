// mlp.hip -- kernels of the learned cost model: a small MLP (features of a
// convolution problem and a tile candidate -> predicted log time), trained on
// the device with plain FP32 kernels: dense forward (+ReLU), mean-squared
// loss gradient, backward through dense+ReLU, weight gradients by
// deterministic block reductions, and Adam.
//
// Shapes: A [N][I] row-major activations, W [I][J], b [J], Z/H [N][J].
#include <hip/hip_runtime.h>

#include "../cg_internal.h"

namespace {

// H[n][j] = act(sum_i A[n][i] W[i][j] + b[j]); relu != 0 applies max(., 0).
__global__ void dense_fwd(const float *__restrict__ A, const float *__restrict__ W, const float *__restrict__ b,
                          float *__restrict__ H, int N, int I, int J, int relu) {
    const int n = blockIdx.x;
    for (int j = threadIdx.x; j < J; j += blockDim.x) {
        if (n >= N) return;
        float s = b[j];
        for (int i = 0; i < I; i++) s += A[(long long)n * I + i] * W[(long long)i * J + j];
        H[(long long)n * J + j] = relu ? fmaxf(s, 0.0f) : s;
    }
}

// dY[n] = 2 (Y[n] - T[n]) / N;  partial[block] = sum over the block of (Y-T)^2
__global__ void mse_grad(const float *__restrict__ Y, const float *__restrict__ T, float *__restrict__ dY,
                         float *__restrict__ partial, int N) {
    __shared__ float red[256];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.0f;
    if (n < N) {
        const float d = Y[n] - T[n];
        dY[n] = 2.0f * d / (float)N;
        e = d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// dA[n][i] = (sum_j dZ[n][j] W[i][j]) * (A[n][i] > 0)   (A is the ReLU output feeding W)
__global__ void dense_bwd_act(const float *__restrict__ dZ, const float *__restrict__ W,
                              const float *__restrict__ A, float *__restrict__ dA, int N, int I, int J) {
    const int n = blockIdx.x;
    if (n >= N) return;
    for (int i = threadIdx.x; i < I; i += blockDim.x) {
        float s = 0.0f;
        for (int j = 0; j < J; j++) s += dZ[(long long)n * J + j] * W[(long long)i * J + j];
        dA[(long long)n * I + i] = A[(long long)n * I + i] > 0.0f ? s : 0.0f;
    }
}

// dW[i][j] = sum_n A[n][i] dZ[n][j]  (block (i, j)),  db[j] = sum_n dZ[n][j]  (block (I, j));
// fixed reduction order, so gradients are deterministic.
__global__ void dense_wgrad(const float *__restrict__ A, const float *__restrict__ dZ, float *__restrict__ dW,
                            float *__restrict__ db, int N, int I, int J) {
    __shared__ float red[256];
    const int i = blockIdx.x, j = blockIdx.y;
    float s = 0.0f;
    for (int n = threadIdx.x; n < N; n += blockDim.x)
        s += (i < I ? A[(long long)n * I + i] : 1.0f) * dZ[(long long)n * J + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (i < I) dW[(long long)i * J + j] = red[0];
        else db[j] = red[0];
    }
}

__global__ void adam_step(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

}  // namespace

extern "C" int cg_mlp_dense_fwd(const float *A, const float *W, const float *b, float *H, int N, int I, int J,
                                int relu, void *stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(dense_fwd, dim3(N), dim3(64), 0, (hipStream_t)stream, A, W, b, H, N, I, J, relu);
    return (int)hipGetLastError();
}

extern "C" int cg_mlp_mse_grad(const float *Y, const float *T, float *dY, float *partial, int N, void *stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(mse_grad, dim3((N + 255) / 256), dim3(256), 0, (hipStream_t)stream, Y, T, dY, partial, N);
    return (int)hipGetLastError();
}

extern "C" int cg_mlp_dense_bwd_act(const float *dZ, const float *W, const float *A, float *dA, int N, int I, int J,
                                    void *stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(dense_bwd_act, dim3(N), dim3(64), 0, (hipStream_t)stream, dZ, W, A, dA, N, I, J);
    return (int)hipGetLastError();
}

extern "C" int cg_mlp_dense_wgrad(const float *A, const float *dZ, float *dW, float *db, int N, int I, int J,
                                  void *stream) {
    hipLaunchKernelGGL(dense_wgrad, dim3(I + 1, J), dim3(256), 0, (hipStream_t)stream, A, dZ, dW, db, N, I, J);
    return (int)hipGetLastError();
}

extern "C" int cg_mlp_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2,
                           float eps, int step, void *stream) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_step, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, b1, b2,
                       eps, c1, c2);
    return (int)hipGetLastError();
}
