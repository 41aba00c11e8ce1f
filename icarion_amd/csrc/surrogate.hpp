// surrogate.hpp - the EHSS scattering surrogate: an MLP 6 -> 128 -> 128 -> 3
// (ReLU), evaluated one wavefront (64 ions) at a time on the matrix cores with
// v_mfma_f32_32x32x8_bf16 (bf16 weights and activations, fp32 accumulation).
//
// Inputs per ion: (u, cos psi, sin psi, f1, f2, f3) - a uniform quantile level,
// the azimuth of the contact normal about the relative velocity, and three
// cluster shape features.  Output: the contact normal in the collision frame
// (t1, t2, v_rel axis), i.e. (sqrt(1-mu^2) cos psi, sqrt(1-mu^2) sin psi, -mu)
// with mu = -n.axis the quantity the reference's offline tables store
// (EHSSCollisionHandler.cpp:253-277).
//
// Layout: the network is computed transposed, H^T = W X^T, so the ions are the
// N dimension of the MFMA.  For 32x32x8 (A: 32x8, B: 8x32, D: 32x32):
//   A lane l: row l%32, k = 4*(l/32) + j   (j = 0..3)
//   B lane l: col l%32, k = 4*(l/32) + j
//   D lane l: col l%32, row = 8*(v/4) + 4*(l/32) + v%4  (v = 0..15)
// so the D fragment of one layer, for k-step s = v/4 of the next, is exactly
// that layer's B fragment: activations chain through registers, no shuffles.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icarion {

constexpr int kMlpIn = 6;
constexpr int kMlpH = 128;
constexpr int kMlpOut = 3;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpOut * kMlpH + kMlpOut;
// natural parameter layout (row-major, out x in)
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpH;

// Packed inference blob (bytes): bf16 A fragments in lane order, then biases.
//   A1: [4 mb][1 ks][64 lanes] x 4 bf16   ( 2048 B)
//   A2: [4 mb][16 ks][64 lanes] x 4 bf16  (32768 B)
//   A3: [1 mb][16 ks][64 lanes] x 4 bf16  ( 8192 B)
//   b1[128], b2[128], b3[32] float         ( 1152 B)
constexpr int kPackA1 = 0;
constexpr int kPackA2 = kPackA1 + 4 * 1 * 64 * 8;
constexpr int kPackA3 = kPackA2 + 4 * 16 * 64 * 8;
constexpr int kPackB1 = kPackA3 + 1 * 16 * 64 * 8;
constexpr int kPackB2 = kPackB1 + 128 * 4;
constexpr int kPackB3 = kPackB2 + 128 * 4;
constexpr int kPackBytes = kPackB3 + 32 * 4;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ short f2bf(float f) {
    // round to nearest even (NaN not expected here)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<short>(u >> 16);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = f2bf(a); r[1] = f2bf(b); r[2] = f2bf(c); r[3] = f2bf(d);
    return r;
}

// Copy the packed blob from global memory into LDS (all threads of the block).
__device__ __forceinline__ void mlp_stage(const uint4* __restrict__ blob, uint4* lds) {
    for (int i = threadIdx.x; i < kPackBytes / 16; i += blockDim.x) lds[i] = blob[i];
    __syncthreads();
}

// Evaluate the MLP for the 64 ions of this wavefront; in[] are this lane's
// inputs (in[6], in[7] = 0), out[] receives this lane's outputs.  Must be
// called by all 64 lanes together (matrix instructions are wave-wide).
__device__ __forceinline__ void mlp_wave_eval(const unsigned char* lds, const float in[8], float out[3]) {
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;  // 0: lanes 0-31, 1: lanes 32-63
    const bf16x4* A1 = reinterpret_cast<const bf16x4*>(lds + kPackA1);
    const bf16x4* A2 = reinterpret_cast<const bf16x4*>(lds + kPackA2);
    const bf16x4* A3 = reinterpret_cast<const bf16x4*>(lds + kPackA3);
    const float* b1 = reinterpret_cast<const float*>(lds + kPackB1);
    const float* b2 = reinterpret_cast<const float*>(lds + kPackB2);
    const float* b3 = reinterpret_cast<const float*>(lds + kPackB3);

    // B operand of layer 1 for the two 32-ion blocks: lane l needs features
    // 4*half..4*half+3 of ion 32*nb + l%32.  Own ion: nb == half.
    float partner[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) partner[j] = __shfl_xor(in[j], 32, 64);
    bf16x4 X[2];
    {
        const float* own = in;
        const float* oth = partner;
        // nb = 0: lanes < 32 own ion (features 0-3), lanes >= 32 partner's ion (4-7)
        X[0] = half == 0 ? pack4(own[0], own[1], own[2], own[3]) : pack4(oth[4], oth[5], oth[6], oth[7]);
        // nb = 1: lanes < 32 partner's ion (0-3), lanes >= 32 own ion (4-7)
        X[1] = half == 0 ? pack4(oth[0], oth[1], oth[2], oth[3]) : pack4(own[4], own[5], own[6], own[7]);
    }
    float res[2][3];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        // layer 1: 4 row blocks of 32 hidden units
        bf16x4 B2[16];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            f32x16 acc = {};
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A1[mb * 64 + lane], X[nb], acc, 0, 0, 0);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int row = 32 * mb + 8 * s + 4 * half;
                const float4 bb = *reinterpret_cast<const float4*>(b1 + row);
                B2[4 * mb + s] = pack4(fmaxf(acc[4 * s] + bb.x, 0.f), fmaxf(acc[4 * s + 1] + bb.y, 0.f),
                                       fmaxf(acc[4 * s + 2] + bb.z, 0.f), fmaxf(acc[4 * s + 3] + bb.w, 0.f));
            }
        }
        // layer 2
        bf16x4 B3[16];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < 16; ++ks)
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A2[(mb * 16 + ks) * 64 + lane], B2[ks], acc, 0, 0, 0);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int row = 32 * mb + 8 * s + 4 * half;
                const float4 bb = *reinterpret_cast<const float4*>(b2 + row);
                B3[4 * mb + s] = pack4(fmaxf(acc[4 * s] + bb.x, 0.f), fmaxf(acc[4 * s + 1] + bb.y, 0.f),
                                       fmaxf(acc[4 * s + 2] + bb.z, 0.f), fmaxf(acc[4 * s + 3] + bb.w, 0.f));
            }
        }
        // layer 3: rows 0..2 of one 32-row block
        f32x16 acc = {};
#pragma unroll
        for (int ks = 0; ks < 16; ++ks)
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A3[ks * 64 + lane], B3[ks], acc, 0, 0, 0);
        // lanes < 32 hold rows 0..3 (v = 0..3) for ion 32*nb + lane
        res[nb][0] = acc[0] + b3[0];
        res[nb][1] = acc[1] + b3[1];
        res[nb][2] = acc[2] + b3[2];
    }
    // ion `lane`: lanes < 32 read their nb = 0 result; lanes >= 32 take the
    // nb = 1 result computed in lane - 32
    float o1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) o1[c] = __shfl_xor(res[1][c], 32, 64);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = half == 0 ? res[0][c] : o1[c];
}

}  // namespace icarion
