// surrogate.hip - training (fp32) and batch inference (bf16 MFMA) kernels of
// the EHSS scattering surrogate (surrogate.hpp).
//
// Training step on a batch of B samples, row-major matrices:
//   Z1 = X W1^T + b1, H1 = relu(Z1), Z2 = H1 W2^T + b2, H2 = relu(Z2),
//   Y = H2 W3^T + b3, loss = sum 0.5 |Y - T|^2 / B, dY = (Y - T) / B,
//   gW3 = dY^T H2, dH2 = dY W3, dZ2 = dH2 * [Z2 > 0], gW2 = dZ2^T H1,
//   dH1 = dZ2 W2, dZ1 = dH1 * [Z1 > 0], gW1 = dZ1^T X, gb = column sums,
// then Adam.  All products go through one fp32 matrix-core GEMM kernel with the
// bias, ReLU, ReLU-derivative and bias-gradient work in its epilogue.
#include <algorithm>

#include "device_common.hpp"
#include "surrogate.hpp"

namespace icarion {

// C[M x N] = op(A)[M x K] * op(B)[K x N] on the matrix cores in fp32
// (v_mfma_f32_32x32x2_f32: fp32 operands and accumulation, the precision of the
// vector fp32 products), row-major storage; op(A) = A (M x K) or A^T (A stored
// K x M), op(B) = B (K x N) or B^T (stored N x K).  A 256-thread block computes a
// 64 x 64 tile, each wavefront one 32 x 32 quarter, over K chunks of 32 staged
// in LDS.  Operand layout of the instruction: A lane l holds row l % 32 at
// k = l / 32, B lane l column l % 32 at k = l / 32; accumulator register v of
// lane l holds row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32.
// blockIdx.z splits K (k_split entries each): split z writes its partial tile
// to W + z M N, and gemm_reduce_kernel sums the splits in order z = 0, 1, ...
// (a fixed order: the result does not depend on the launch).  Epilogues on the
// whole product: kEpiStore (C = acc, or += with accumulate), kEpiBias (Z = C =
// acc + bias[n]; H = relu(Z) when H is given), kEpiMask (C = acc where mask > 0,
// else 0: the ReLU derivative).  With colsum given (TA only), the blocks of
// column tile 0 also sum op(A) over k: colsum[m] = sum_k A[k][m], the bias
// gradient of a weight gradient dZ^T H (partials at Wsum + z M when split).
constexpr int kEpiStore = 0;
constexpr int kEpiBias = 1;
constexpr int kEpiMask = 2;

typedef float f32x16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void gemm_epilogue(int mode, float v, size_t idx, int n, float* C, const float* bias,
                                              float* H, const float* mask, int accumulate) {
    if (mode == kEpiBias) {
        const float z = v + bias[n];
        C[idx] = z;
        if (H) H[idx] = fmaxf(z, 0.f);
    } else if (mode == kEpiMask) {
        C[idx] = mask[idx] > 0.f ? v : 0.f;
    } else {
        C[idx] = accumulate ? C[idx] + v : v;
    }
}

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_mfma_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                        float* C, int M, int N, int K, int k_split, int mode,
                                                        const float* __restrict__ bias, float* H,
                                                        const float* __restrict__ mask, int accumulate,
                                                        float* __restrict__ W, float* colsum,
                                                        float* __restrict__ Wsum) {
    __shared__ float As[32][64 + 4];
    __shared__ float Bs[32][64 + 4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int kb = blockIdx.z * k_split;
    const int ke = min(K, kb + k_split);
    const bool sums = TA && colsum != nullptr && blockIdx.x == 0 && wave == 0;
    f32x16_t acc = {};
    float csum = 0.f;
    // Each thread moves 8 elements of each operand per K chunk; the next chunk's
    // loads are issued before the current chunk's matrix instructions.
    float ra[8], rb[8];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = static_cast<int>(threadIdx.x) + 256 * i;
            {
                const int kk = TA ? e >> 6 : e & 31;
                const int mm = TA ? e & 63 : e >> 5;
                const int gk = k0 + kk, gm = m0 + mm;
                ra[i] = gk < ke && gm < M
                            ? (TA ? A[static_cast<size_t>(gk) * M + gm] : A[static_cast<size_t>(gm) * K + gk])
                            : 0.f;
            }
            {
                const int kk = TB ? e & 31 : e >> 6;
                const int nn = TB ? e >> 5 : e & 63;
                const int gk = k0 + kk, gn = n0 + nn;
                rb[i] = gk < ke && gn < N
                            ? (TB ? B[static_cast<size_t>(gn) * K + gk] : B[static_cast<size_t>(gk) * N + gn])
                            : 0.f;
            }
        }
    };
    if (kb < ke) load(kb);
    for (int k0 = kb; k0 < ke; k0 += 32) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = static_cast<int>(threadIdx.x) + 256 * i;
            As[TA ? e >> 6 : e & 31][TA ? e & 63 : e >> 5] = ra[i];
            Bs[TB ? e & 31 : e >> 6][TB ? e >> 5 : e & 63] = rb[i];
        }
        __syncthreads();
        if (k0 + 32 < ke) load(k0 + 32);
#pragma unroll
        for (int kk = 0; kk < 32; kk += 2) {
            const float a = As[kk + (lane >> 5)][wm + (lane & 31)];
            const float b = Bs[kk + (lane >> 5)][wn + (lane & 31)];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
        }
        if (sums) {
#pragma unroll 8
            for (int kk = 0; kk < 32; ++kk) csum += As[kk][lane];
        }
        __syncthreads();
    }
    const bool split = gridDim.z > 1;
    if (sums && m0 + lane < M) {
        if (split)
            Wsum[static_cast<size_t>(blockIdx.z) * M + m0 + lane] = csum;
        else
            colsum[m0 + lane] = csum;
    }
    const int col = n0 + wn + (lane & 31);
    float* part = split ? W + static_cast<size_t>(blockIdx.z) * M * N : nullptr;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = m0 + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (row < M && col < N) {
            const size_t idx = static_cast<size_t>(row) * N + col;
            if (split)
                part[idx] = acc[v];
            else
                gemm_epilogue(mode, acc[v], idx, col, C, bias, H, mask, accumulate);
        }
    }
}

// C = the epilogue of the sum over the splits of W[z][M x N], and colsum[m] =
// the sum over the splits of Wsum[z][m], both summed in the order z = 0, 1, ...
__global__ void gemm_reduce_kernel(const float* __restrict__ W, int splits, int M, int N, float* C, int mode,
                                   const float* __restrict__ bias, float* H, const float* __restrict__ mask,
                                   int accumulate, const float* __restrict__ Wsum, float* colsum) {
    const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t mn = static_cast<size_t>(M) * N;
    if (colsum && idx < static_cast<size_t>(M)) {
        float c = 0.f;
#pragma unroll 8
        for (int z = 0; z < splits; ++z) c += Wsum[z * static_cast<size_t>(M) + idx];
        colsum[idx] = c;
    }
    if (idx >= mn) return;
    float v = 0.f;
#pragma unroll 8
    for (int z = 0; z < splits; ++z) v += W[z * mn + idx];
    gemm_epilogue(mode, v, idx, static_cast<int>(idx % N), C, bias, H, mask, accumulate);
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

struct GemmEpi {
    int mode = kEpiStore;
    const float* bias = nullptr;
    float* H = nullptr;
    const float* mask = nullptr;
    bool accumulate = false;
    float* colsum = nullptr;  // bias gradient (op(A) = A^T only)
};

// C = op(A) op(B) with the epilogue e.  A product whose K (the batch) exceeds
// 256 with op(A) = A^T - a weight gradient, M, N <= kMlpH - is split over K
// into at most kGemmSplitMax ranges of a multiple of 32, partials in ws
// (kGemmWorkspaceFloats).
hipError_t gemm_f32(bool ta, bool tb, const float* A, const float* B, float* C, int M, int N, int K,
                    const GemmEpi& e, float* ws, hipStream_t s) {
    if (e.colsum && !ta) return hipErrorInvalidValue;
    int splits = 1, k_split = K;
    if (ta && K > 256) {
        if (M > kMlpH || N > kMlpH || !ws) return hipErrorInvalidValue;
        splits = std::min(kGemmSplitMax, (K + 255) / 256);
        k_split = ((K + splits - 1) / splits + 31) / 32 * 32;
        splits = (K + k_split - 1) / k_split;
    }
    float* W = splits > 1 ? ws : nullptr;
    float* Wsum = splits > 1 ? ws + static_cast<size_t>(kGemmSplitMax) * kMlpH * kMlpH : nullptr;
    const dim3 grid((N + 63) / 64, (M + 63) / 64, splits);
    const int acc = e.accumulate ? 1 : 0;
#define ICARION_GEMM_LAUNCH(TA_, TB_)                                                                        \
    hipLaunchKernelGGL((gemm_mfma_kernel<TA_, TB_>), grid, dim3(256), 0, s, A, B, C, M, N, K, k_split, e.mode, \
                       e.bias, e.H, e.mask, acc, W, e.colsum, Wsum)
    if (!ta && !tb) ICARION_GEMM_LAUNCH(false, false);
    else if (!ta && tb) ICARION_GEMM_LAUNCH(false, true);
    else if (ta && !tb) ICARION_GEMM_LAUNCH(true, false);
    else ICARION_GEMM_LAUNCH(true, true);
#undef ICARION_GEMM_LAUNCH
    if (splits > 1) {
        const size_t mn = static_cast<size_t>(M) * N;
        hipLaunchKernelGGL(gemm_reduce_kernel, dim3(static_cast<uint32_t>((mn + 255) / 256)), dim3(256), 0, s, W,
                           splits, M, N, C, e.mode, e.bias, e.H, e.mask, acc, Wsum, e.colsum);
    }
    return hipGetLastError();
}

struct TrainBuffers {
    float *X, *T;            // batch views
    float *Z1, *H1, *Z2, *H2, *Y, *dY, *dH2, *dH1;
    float *params, *grad, *m, *v, *loss;
    float* ws;               // kGemmWorkspaceFloats
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
    auto bias = [](const float* b, float* H) {
        GemmEpi g;
        g.mode = kEpiBias; g.bias = b; g.H = H;
        return g;
    };
    auto masked = [](const float* Z) {
        GemmEpi g;
        g.mode = kEpiMask; g.mask = Z;
        return g;
    };
    auto with_colsum = [](float* gb) {
        GemmEpi g;
        g.colsum = gb;
        return g;
    };
    // forward: Z1 = X W1^T + b1, H1 = relu(Z1); Z2, H2 likewise; Y = H2 W3^T + b3
    if ((e = gemm_f32(false, true, t.X, W1, t.Z1, B, kMlpH, kMlpIn, bias(b1, t.H1), t.ws, s))) return e;
    if ((e = gemm_f32(false, true, t.H1, W2, t.Z2, B, kMlpH, kMlpH, bias(b2, t.H2), t.ws, s))) return e;
    if ((e = gemm_f32(false, true, t.H2, W3, t.Y, B, kMlpOut, kMlpH, bias(b3, nullptr), t.ws, s))) return e;
    const int th = 256;
    hipLaunchKernelGGL(out_grad_kernel, dim3((B * kMlpOut + th - 1) / th), dim3(th), 0, s, t.Y, t.T, t.dY,
                       B * kMlpOut, 1.0f / static_cast<float>(B), t.loss);
    // backward: dH2 holds dZ2 = (dY W3) [Z2 > 0], dH1 holds dZ1 = (dZ2 W2) [Z1 > 0]
    if ((e = gemm_f32(true, false, t.dY, t.H2, gW3, kMlpOut, kMlpH, B, with_colsum(gb3), t.ws, s))) return e;
    if ((e = gemm_f32(false, false, t.dY, W3, t.dH2, B, kMlpH, kMlpOut, masked(t.Z2), t.ws, s))) return e;
    if ((e = gemm_f32(true, false, t.dH2, t.H1, gW2, kMlpH, kMlpH, B, with_colsum(gb2), t.ws, s))) return e;
    if ((e = gemm_f32(false, false, t.dH2, W2, t.dH1, B, kMlpH, kMlpH, masked(t.Z1), t.ws, s))) return e;
    if ((e = gemm_f32(true, false, t.dH1, t.X, gW1, kMlpH, kMlpIn, B, with_colsum(gb1), t.ws, s))) return e;
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
