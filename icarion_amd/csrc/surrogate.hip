// surrogate.hip - training (fp32) and batch inference (bf16 MFMA) kernels of
// the EHSS scattering surrogate (surrogate.hpp).
//
// Training step on a batch of B samples, row-major matrices:
//   Z1 = X W1^T + b1, H1 = relu(Z1), Z2 = H1 W2^T + b2, H2 = relu(Z2),
//   Y = H2 W3^T + b3, loss = sum 0.5 |Y - T|^2 / B, dY = (Y - T) / B,
//   gW3 = dY^T H2, dH2 = dY W3, dZ2 = dH2 * [Z2 > 0], gW2 = dZ2^T H1,
//   dH1 = dZ2 W2, dZ1 = dH1 * [Z1 > 0], gW1 = dZ1^T X, gb = column sums,
// then Adam.  All products go through one LDS-tiled fp32 GEMM kernel.
#include "device_common.hpp"
#include "surrogate.hpp"

namespace icarion {

// C[M x N] = op(A)[M x K] * op(B)[K x N] (+ C if accumulate); row-major storage,
// op(A) = A (M x K) or A^T (A stored K x M), op(B) = B (K x N) or B^T (stored N x K).
// 64 x 64 output tile per 256-thread block, 4 x 4 outputs per thread, K tile 16.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_f32_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                       float* __restrict__ Cm, int M, int N, int K, int accumulate) {
    __shared__ float As[16][64 + 1];
    __shared__ float Bs[16][64 + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            const int kk = e / 64, mm = e % 64;
            const int gm = m0 + mm, gk = k0 + kk;
            float va = 0.f;
            if (gm < M && gk < K) va = TA ? A[static_cast<size_t>(gk) * M + gm] : A[static_cast<size_t>(gm) * K + gk];
            As[kk][mm] = va;
            const int gn = n0 + mm;
            float vb = 0.f;
            if (gn < N && gk < K) vb = TB ? B[static_cast<size_t>(gn) * K + gk] : B[static_cast<size_t>(gk) * N + gn];
            Bs[kk][mm] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gm = m0 + ty * 4 + i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gn = n0 + tx * 4 + j;
            if (gn >= N) continue;
            const size_t idx = static_cast<size_t>(gm) * N + gn;
            Cm[idx] = accumulate ? Cm[idx] + acc[i][j] : acc[i][j];
        }
    }
}

// H = relu(Z + b) (Z, H: rows x cols), keeps Z for the backward pass
__global__ void bias_relu_kernel(float* Z, float* H, const float* b, int rows, int cols, int relu) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(rows) * cols) return;
    const float z = Z[i] + b[i % cols];
    Z[i] = z;
    H[i] = relu ? fmaxf(z, 0.f) : z;
}

// dY = (Y - T) / B; loss accumulated (sum of 0.5 d^2 / B)
__global__ void out_grad_kernel(const float* Y, const float* T, float* dY, int n, float inv_b, float* loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float d = Y[i] - T[i];
        dY[i] = d * inv_b;
        l = 0.5f * d * d * inv_b;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

__global__ void relu_bwd_kernel(float* dH, const float* Z, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && !(Z[i] > 0.f)) dH[i] = 0.f;
}

// g[c] = sum_r D[r][c]; one block per column
__global__ void colsum_kernel(const float* D, float* g, int rows, int cols) {
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) s += D[static_cast<size_t>(r) * cols + c];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) g[c] = red[0];
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, int n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// Batch inference: each wavefront evaluates 64 samples on the matrix cores.
__global__ void __launch_bounds__(256) mlp_infer_kernel(const uint4* __restrict__ blob, const float* __restrict__ X,
                                                        float* __restrict__ Y, uint32_t n) {
    __shared__ uint4 lds[kPackBytes / 16];
    mlp_stage(blob, lds);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float in[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
#pragma unroll
        for (int j = 0; j < kMlpIn; ++j) in[j] = X[static_cast<size_t>(i) * kMlpIn + j];
    }
    float out[3];
    mlp_wave_eval(reinterpret_cast<const unsigned char*>(lds), in, out);
    if (i < n) {
        Y[static_cast<size_t>(i) * 3 + 0] = out[0];
        Y[static_cast<size_t>(i) * 3 + 1] = out[1];
        Y[static_cast<size_t>(i) * 3 + 2] = out[2];
    }
}

// ---- host-side launchers -----------------------------------------------------

hipError_t gemm_f32(bool ta, bool tb, const float* A, const float* B, float* C, int M, int N, int K,
                    bool accumulate, hipStream_t s) {
    const dim3 grid((N + 63) / 64, (M + 63) / 64);
    if (!ta && !tb) hipLaunchKernelGGL((gemm_f32_kernel<false, false>), grid, dim3(256), 0, s, A, B, C, M, N, K, accumulate ? 1 : 0);
    else if (!ta && tb) hipLaunchKernelGGL((gemm_f32_kernel<false, true>), grid, dim3(256), 0, s, A, B, C, M, N, K, accumulate ? 1 : 0);
    else if (ta && !tb) hipLaunchKernelGGL((gemm_f32_kernel<true, false>), grid, dim3(256), 0, s, A, B, C, M, N, K, accumulate ? 1 : 0);
    else hipLaunchKernelGGL((gemm_f32_kernel<true, true>), grid, dim3(256), 0, s, A, B, C, M, N, K, accumulate ? 1 : 0);
    return hipGetLastError();
}

struct TrainBuffers {
    float *X, *T;            // batch views
    float *Z1, *H1, *Z2, *H2, *Y, *dY, *dH2, *dH1;
    float *params, *grad, *m, *v, *loss;
    int B;
};

// Forward + backward of one batch; grad = d(loss)/d(params); loss += batch loss
hipError_t mlp_fwd_bwd(const TrainBuffers& t, hipStream_t s) {
    const int B = t.B;
    const float* W1 = t.params + kOffW1; const float* b1 = t.params + kOffB1;
    const float* W2 = t.params + kOffW2; const float* b2 = t.params + kOffB2;
    const float* W3 = t.params + kOffW3; const float* b3 = t.params + kOffB3;
    float* gW1 = t.grad + kOffW1; float* gb1 = t.grad + kOffB1;
    float* gW2 = t.grad + kOffW2; float* gb2 = t.grad + kOffB2;
    float* gW3 = t.grad + kOffW3; float* gb3 = t.grad + kOffB3;
    hipError_t e;
    const int th = 256;
    auto blocks = [&](size_t n) { return dim3(static_cast<uint32_t>((n + th - 1) / th)); };
    if ((e = gemm_f32(false, true, t.X, W1, t.Z1, B, kMlpH, kMlpIn, false, s))) return e;
    hipLaunchKernelGGL(bias_relu_kernel, blocks(static_cast<size_t>(B) * kMlpH), dim3(th), 0, s, t.Z1, t.H1, b1, B, kMlpH, 1);
    if ((e = gemm_f32(false, true, t.H1, W2, t.Z2, B, kMlpH, kMlpH, false, s))) return e;
    hipLaunchKernelGGL(bias_relu_kernel, blocks(static_cast<size_t>(B) * kMlpH), dim3(th), 0, s, t.Z2, t.H2, b2, B, kMlpH, 1);
    if ((e = gemm_f32(false, true, t.H2, W3, t.Y, B, kMlpOut, kMlpH, false, s))) return e;
    hipLaunchKernelGGL(bias_relu_kernel, blocks(static_cast<size_t>(B) * kMlpOut), dim3(th), 0, s, t.Y, t.Y, b3, B, kMlpOut, 0);
    hipLaunchKernelGGL(out_grad_kernel, blocks(static_cast<size_t>(B) * kMlpOut), dim3(th), 0, s, t.Y, t.T, t.dY,
                       B * kMlpOut, 1.0f / static_cast<float>(B), t.loss);
    if ((e = gemm_f32(true, false, t.dY, t.H2, gW3, kMlpOut, kMlpH, B, false, s))) return e;
    hipLaunchKernelGGL(colsum_kernel, dim3(kMlpOut), dim3(256), 0, s, t.dY, gb3, B, kMlpOut);
    if ((e = gemm_f32(false, false, t.dY, W3, t.dH2, B, kMlpH, kMlpOut, false, s))) return e;
    hipLaunchKernelGGL(relu_bwd_kernel, blocks(static_cast<size_t>(B) * kMlpH), dim3(th), 0, s, t.dH2, t.Z2, static_cast<size_t>(B) * kMlpH);
    if ((e = gemm_f32(true, false, t.dH2, t.H1, gW2, kMlpH, kMlpH, B, false, s))) return e;
    hipLaunchKernelGGL(colsum_kernel, dim3(kMlpH), dim3(256), 0, s, t.dH2, gb2, B, kMlpH);
    if ((e = gemm_f32(false, false, t.dH2, W2, t.dH1, B, kMlpH, kMlpH, false, s))) return e;
    hipLaunchKernelGGL(relu_bwd_kernel, blocks(static_cast<size_t>(B) * kMlpH), dim3(th), 0, s, t.dH1, t.Z1, static_cast<size_t>(B) * kMlpH);
    if ((e = gemm_f32(true, false, t.dH1, t.X, gW1, kMlpH, kMlpIn, B, false, s))) return e;
    hipLaunchKernelGGL(colsum_kernel, dim3(kMlpH), dim3(256), 0, s, t.dH1, gb1, B, kMlpH);
    return hipGetLastError();
}

hipError_t adam_step(const TrainBuffers& t, float lr, float b1, float b2, float eps, int step, hipStream_t s) {
    const float c1 = 1.f / (1.f - powf(b1, static_cast<float>(step)));
    const float c2 = 1.f / (1.f - powf(b2, static_cast<float>(step)));
    hipLaunchKernelGGL(adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, t.params, t.grad, t.m, t.v,
                       kMlpParams, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t mlp_infer(const uint4* blob, const float* X, float* Y, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((n + 255) / 256), dim3(256), 0, s, blob, X, Y, n);
    return hipGetLastError();
}

}  // namespace icarion
