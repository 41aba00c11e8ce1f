// surrogate.hpp - the EHSS scattering surrogate: an MLP 6 -> 128 -> 128 -> 3
// (ReLU), evaluated one wavefront (64 ions) at a time on the matrix cores
// (bf16 weights and activations, fp32 accumulation).
//
// Inputs per ion: (u, cos psi, sin psi, f1, f2, f3) - a uniform quantile level,
// the azimuth of the contact normal about the relative velocity, and three
// cluster shape features.  Output: the contact normal in the collision frame
// (t1, t2, v_rel axis), i.e. (sqrt(1-mu^2) cos psi, sqrt(1-mu^2) sin psi, -mu)
// with mu = -n.axis the quantity the reference's offline tables store
// (EHSSCollisionHandler.cpp:253-277).
//
// Layout: the network is computed transposed, H^T = W X^T, so the ions are the
// N dimension of the matrix instructions (two 32-ion blocks per wavefront).
// Every 32x32 output block has the accumulator layout
//   D lane l: col l%32, row = 8*(v/4) + 4*(l/32) + v%4  (v = 0..15).
// Layer 1 uses v_mfma_f32_32x32x8_bf16 (K = 8: the six inputs, a constant 1
// that carries the bias, a zero):
//   A lane l: row l%32, k = 4*(l/32) + j;  B lane l: col l%32, k = 4*(l/32) + j.
// Layers 2 and 3 use v_mfma_f32_32x32x16_bf16 (K = 16 per instruction; layer
// 2's bias is the accumulators' initial value, in fp32): an
// operand slot (half h = l/32, element j = 0..7) holds the same k for A and B,
// and the sum over k does not depend on which hidden unit a slot carries.  The
// B operand of K-chunk c = 2*mb + q is taken straight from accumulator
// registers v = 8q + j of the previous layer's row block mb, so slot (h, j)
// carries hidden unit 32*mb + 16*q + 8*(j/4) + 4*h + j%4; the host packs the
// weight columns in that order.  Activations chain through registers, no
// shuffles.  Layer 3 is consumed block by block inside layer 2's loop, so
// layer 2's activations never exist all at once.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace icarion {

constexpr int kMlpIn = 6;
constexpr int kMlpH = 128;
constexpr int kMlpOut = 3;
// Workgroup size of the single-cloud surrogate step kernel, which runs three
// wavefronts per SIMD (3 workgroups of 50 KB LDS per CU).  Eight-wavefront
// workgroups at four wavefronts per SIMD spill registers and run no faster.
constexpr int kSurrogateBlock = 256;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpOut * kMlpH + kMlpOut;
// natural parameter layout (row-major, out x in)
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpH;
// Training GEMMs over the batch (K = batch rows) are split into at most
// kGemmSplitMax K ranges; the workspace holds their partial products
// ([split][M][N], M, N <= kMlpH) and partial column sums ([split][M]).
constexpr int kGemmSplitMax = 64;
constexpr int kGemmWorkspaceFloats = kGemmSplitMax * (kMlpH * kMlpH + kMlpH);

// Packed inference blob (bytes): bf16 A fragments in lane order, then biases.
//   A1: [4 mb][64 lanes] x 4 bf16         ( 2048 B)  column 6 = b1
//   A2: [4 mb][8 kc][64 lanes] x 8 bf16   (32768 B)
//   A3: [8 kc][64 lanes] x 8 bf16         ( 8192 B)  rows 3..31 zero
//   b2[128] float                         (  512 B)  layer 2's accumulators start at b2
//   b3[32] float                          (  128 B)
constexpr int kPackA1 = 0;
constexpr int kPackA2 = kPackA1 + 4 * 64 * 8;
constexpr int kPackA3 = kPackA2 + 4 * 8 * 64 * 16;
constexpr int kPackB2 = kPackA3 + 8 * 64 * 16;
constexpr int kPackB3 = kPackB2 + kMlpH * 4;
constexpr int kPackBytes = kPackB3 + 32 * 4;
static_assert(kPackBytes % 16 == 0, "the blob is staged as uint4");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32v4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Hidden unit carried by slot (half h, element j) of K-chunk kc (layers 2, 3).
__host__ __device__ constexpr int mlp_chunk_unit(int kc, int h, int j) {
    return 16 * kc + 8 * (j / 4) + 4 * h + j % 4;
}

// float -> bf16 bits, round to nearest even (host packing; the device converts
// with v_cvt_pk_bf16_f32, which rounds the same way).
inline uint16_t mlp_bf16_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

// Pack fp32 parameters (natural layout above) into the blob.
inline void mlp_pack_blob(const float* p, unsigned char* blob) {
    memset(blob, 0, kPackBytes);
    auto put = [&](int byte_off, int j, float v) {
        const uint16_t b = mlp_bf16_bits(v);
        memcpy(blob + byte_off + 2 * j, &b, 2);
    };
    const float* W1 = p + kOffW1;
    const float* b1 = p + kOffB1;
    const float* W2 = p + kOffW2;
    const float* W3 = p + kOffW3;
    for (int mb = 0; mb < 4; ++mb)
        for (int lane = 0; lane < 64; ++lane)
            for (int j = 0; j < 4; ++j) {
                const int row = 32 * mb + lane % 32, k = 4 * (lane / 32) + j;
                const float v = k < kMlpIn ? W1[row * kMlpIn + k] : (k == kMlpIn ? b1[row] : 0.f);
                put(kPackA1 + (mb * 64 + lane) * 8, j, v);
            }
    for (int mb = 0; mb < 4; ++mb)
        for (int kc = 0; kc < 8; ++kc)
            for (int lane = 0; lane < 64; ++lane)
                for (int j = 0; j < 8; ++j) {
                    const int row = 32 * mb + lane % 32;
                    put(kPackA2 + ((mb * 8 + kc) * 64 + lane) * 16, j,
                        W2[row * kMlpH + mlp_chunk_unit(kc, lane / 32, j)]);
                }
    memcpy(blob + kPackB2, p + kOffB2, kMlpH * 4);
    for (int kc = 0; kc < 8; ++kc)
        for (int lane = 0; lane < 64; ++lane)
            for (int j = 0; j < 8; ++j) {
                const int row = lane % 32, k = mlp_chunk_unit(kc, lane / 32, j);
                put(kPackA3 + (kc * 64 + lane) * 16, j, row < kMlpOut ? W3[row * kMlpH + k] : 0.f);
            }
    memcpy(blob + kPackB3, p + kOffB3, kMlpOut * 4);
}

// Two floats -> a bf16 pair (round to nearest even), then ReLU on the pair as
// two signed 16-bit integers: a negative bf16 has its sign bit set, so the max
// with 0 gives +0.  Rounding keeps the sign, so this is bf16(relu(x)).
__device__ __forceinline__ uint32_t relu_pk(float a, float b) {
    const f32x2 v = {a, b};
    const bf16x2 r = __builtin_convertvector(v, bf16x2);
    const s16x2 m = __builtin_elementwise_max(__builtin_bit_cast(s16x2, r), (s16x2){0, 0});
    return __builtin_bit_cast(uint32_t, m);
}

// bf16 ReLU of accumulator registers c[base .. base+7] as one B operand.
__device__ __forceinline__ bf16x8 relu_pack8(const f32x16& c, int base) {
    const u32v4 u = {relu_pk(c[base], c[base + 1]), relu_pk(c[base + 2], c[base + 3]),
                     relu_pk(c[base + 4], c[base + 5]), relu_pk(c[base + 6], c[base + 7])};
    return __builtin_bit_cast(bf16x8, u);
}

// Copy the packed blob from global memory into LDS (all threads of the block).
__device__ __forceinline__ void mlp_stage(const uint4* __restrict__ blob, uint4* lds) {
    for (int i = threadIdx.x; i < kPackBytes / 16; i += blockDim.x) lds[i] = blob[i];
    __syncthreads();
}

// bf16 of four floats (round to nearest even), as one layer-1 operand slot quad.
__device__ __forceinline__ bf16x4 mlp_pack4(float a, float b, float c, float d) {
    bf16v4 r;
    r[0] = (__bf16)a; r[1] = (__bf16)b; r[2] = (__bf16)c; r[3] = (__bf16)d;
    return __builtin_bit_cast(bf16x4, r);
}

// The MLP for 64 ions on the matrix cores.  X0 / X1 are layer 1's B operands
// of the two 32-ion blocks: lane l carries slots 4*(l/32) .. 4*(l/32)+3 of
// ion 32*nb + l%32, slots = (x0..x5, 1, 0) in bf16.  On return lanes < 32 hold
// the outputs of ion lane in r0 and of ion 32 + lane in r1 (lanes >= 32 hold
// nothing useful).  Must be called by all 64 lanes together (matrix
// instructions are wave-wide).
__device__ __forceinline__ void mlp_wave_core(const unsigned char* lds, bf16x4 X0, bf16x4 X1,
                                              float r0[3], float r1[3]) {
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;  // 0: lanes 0-31, 1: lanes 32-63
    const bf16x4* A1 = reinterpret_cast<const bf16x4*>(lds + kPackA1);
    const bf16x8* A2 = reinterpret_cast<const bf16x8*>(lds + kPackA2);
    const bf16x8* A3 = reinterpret_cast<const bf16x8*>(lds + kPackA3);
    const float* b2 = reinterpret_cast<const float*>(lds + kPackB2);
    const float* b3 = reinterpret_cast<const float*>(lds + kPackB3);
    // Both 32-ion blocks go through each weight fragment together, so every
    // A fragment is read from LDS once per wavefront.
    // layer 1 (bias in slot 6): chunk kc = 2*mb + q holds rows 32*mb + 16*q + ...
    bf16x8 H0[8], H1[8];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const bf16x4 a = A1[mb * 64 + lane];
        f32x16 c0 = {}, c1 = {};
        c0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, X0, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, X1, c1, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            H0[2 * mb + q] = relu_pack8(c0, 8 * q);
            H1[2 * mb + q] = relu_pack8(c1, 8 * q);
        }
    }
    // layers 2 and 3, one 32-unit block of layer 2 at a time
    f32x16 y0 = {}, y1 = {};
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const bf16x8* A2m = A2 + mb * 8 * 64 + lane;
        // the accumulators start at the bias of their rows:
        // register v holds row 32 mb + 8 (v / 4) + 4 half + v % 4
        f32x16 c0;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const float4 bq = *reinterpret_cast<const float4*>(b2 + 32 * mb + 8 * g + 4 * half);
            c0[4 * g] = bq.x;
            c0[4 * g + 1] = bq.y;
            c0[4 * g + 2] = bq.z;
            c0[4 * g + 3] = bq.w;
        }
        f32x16 c1 = c0;
#pragma unroll
        for (int kc = 0; kc < 8; ++kc) {
            const bf16x8 a = A2m[kc * 64];
            c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, H0[kc], c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, H1[kc], c1, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const bf16x8 a3 = A3[(2 * mb + q) * 64 + lane];
            y0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a3, relu_pack8(c0, 8 * q), y0, 0, 0, 0);
            y1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a3, relu_pack8(c1, 8 * q), y1, 0, 0, 0);
        }
    }
    // lanes < 32 hold output rows 0..2 (v = 0..2) of ion 32*nb + lane in y{nb}
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r0[c] = y0[c] + b3[c];
        r1[c] = y1[c] + b3[c];
    }
}

// The MLP for one 32-ion block: X is layer 1's B operand (lane l carries slots
// 4*(l/32) .. 4*(l/32)+3 of ion l%32); on return lanes < 32 hold the outputs of
// ion lane in r.  Half the live registers of mlp_wave_core (each weight
// fragment is read from LDS once per block instead of once per two blocks);
// every output equals mlp_wave_core's for the same operand.
__device__ __forceinline__ void mlp_block_core(const unsigned char* lds, bf16x4 X, float r[3]) {
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    const bf16x4* A1 = reinterpret_cast<const bf16x4*>(lds + kPackA1);
    const bf16x8* A2 = reinterpret_cast<const bf16x8*>(lds + kPackA2);
    const bf16x8* A3 = reinterpret_cast<const bf16x8*>(lds + kPackA3);
    const float* b2 = reinterpret_cast<const float*>(lds + kPackB2);
    const float* b3 = reinterpret_cast<const float*>(lds + kPackB3);
    bf16x8 H[8];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        f32x16 c = {};
        c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A1[mb * 64 + lane], X, c, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 2; ++q) H[2 * mb + q] = relu_pack8(c, 8 * q);
    }
    f32x16 y = {};
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const bf16x8* A2m = A2 + mb * 8 * 64 + lane;
        f32x16 c;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const float4 bq = *reinterpret_cast<const float4*>(b2 + 32 * mb + 8 * g + 4 * half);
            c[4 * g] = bq.x;
            c[4 * g + 1] = bq.y;
            c[4 * g + 2] = bq.z;
            c[4 * g + 3] = bq.w;
        }
#pragma unroll
        for (int kc = 0; kc < 8; ++kc) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A2m[kc * 64], H[kc], c, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 2; ++q)
            y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A3[(2 * mb + q) * 64 + lane], relu_pack8(c, 8 * q), y, 0, 0, 0);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = y[c] + b3[c];
}

// Evaluate the MLP for the 64 ions of this wavefront; in[0..5] are this lane's
// inputs, out[] receives this lane's outputs.  Must be called by all 64 lanes
// together.
__device__ __forceinline__ void mlp_wave_eval(const unsigned char* lds, const float in[6], float out[3]) {
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    // B operand of layer 1 for the two 32-ion blocks: lane l needs slots
    // 4*half..4*half+3 of ion 32*nb + l%32, slots = (x0..x5, 1, 0).
    float partner[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) partner[j] = __shfl_xor(in[j], 32, 64);
    // nb = 0: lanes < 32 own ion (slots 0-3), lanes >= 32 partner's ion (4-7);
    // nb = 1: lanes < 32 partner's ion (0-3), lanes >= 32 own ion (4-7)
    const bf16x4 X0 = half == 0 ? mlp_pack4(in[0], in[1], in[2], in[3])
                                : mlp_pack4(partner[4], partner[5], 1.0f, 0.0f);
    const bf16x4 X1 = half == 0 ? mlp_pack4(partner[0], partner[1], partner[2], partner[3])
                                : mlp_pack4(in[4], in[5], 1.0f, 0.0f);
    float res0[3], res1[3];
    mlp_wave_core(lds, X0, X1, res0, res1);
    // ion `lane`: lanes < 32 read their nb = 0 result; lanes >= 32 take the
    // nb = 1 result computed in lane - 32
    float o1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) o1[c] = __shfl_xor(res1[c], 32, 64);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = half == 0 ? res0[c] : o1[c];
}

// The MLP for the colliding lanes of this wavefront only (no barriers), both
// 32-lane halves together through mlp_wave_core.  `coll` lanes receive their
// outputs in out[].  Returns false when no lane of the wavefront collides.
// Must be called by all 64 lanes together.
__device__ __forceinline__ bool mlp_wave_eval_any(const unsigned char* lds, bool coll, const float in[6],
                                                  float out[3]) {
    if (!__any(coll)) return false;
    // a wave in its matrix-core section runs at raised priority so that its
    // MFMA chain keeps issuing while the other wave of the SIMD fills the gaps
    // with vector work
    __builtin_amdgcn_s_setprio(1);
    mlp_wave_eval(lds, in, out);
    __builtin_amdgcn_s_setprio(0);
    return true;
}

// LDS exchange of a workgroup (up to 4 wavefronts) whose colliding ions are
// gathered into full 64-ion batches for the MLP.  Wavefront w writes the inputs
// of its colliding lanes, in lane order, to slots 64 w .. 64 w + count - 1, as
// the eight bf16 operand slots of layer 1 (x0..x5, 1, 0); the counts are
// double-buffered by step parity, since a workgroup without any colliding ion
// passes only one barrier per step.
template <int WAVES>
struct MlpExchange {
    static constexpr int kWaves = WAVES;
    u32v4 in[64 * WAVES];
    float out[kMlpOut][64 * WAVES];
    uint32_t count[2][WAVES];
};

// Zero the counts (wavefronts beyond blockDim.x never write theirs); the
// barrier that follows in mlp_stage publishes them.
template <int WAVES>
__device__ __forceinline__ void mlp_exchange_init(MlpExchange<WAVES>& X) {
    if (threadIdx.x < 2 * WAVES) X.count[threadIdx.x / WAVES][threadIdx.x % WAVES] = 0u;
}

// MLP outputs for the colliding lanes of the whole workgroup.  Every lane of
// every wavefront calls this once per step (it holds barriers); `coll` lanes
// pass their inputs in in[] and receive their outputs in out[].  The colliding
// ions are numbered wavefront by wavefront, lane by lane; batch b holds ions
// 64 b .. 64 b + 63 of that numbering and is evaluated by wavefront
// (b + rot) mod (waves), so that successive steps spread the matrix work over
// the SIMDs.  Batches are padded with zero operands whose outputs are dropped.
// An ion's outputs do not depend on its batch or column: each output column of
// a matrix instruction is its own dot product, and the operands are the bf16
// values mlp_wave_eval forms.  Returns false when no ion of the workgroup
// collides (out[] is then untouched).
template <int WAVES>
__device__ __forceinline__ bool mlp_group_eval(const unsigned char* lds, MlpExchange<WAVES>& X, uint32_t step,
                                               bool coll, const float in[kMlpIn], float out[kMlpOut],
                                               uint32_t rot) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t nw = __builtin_amdgcn_readfirstlane(blockDim.x >> 6);
    const unsigned long long m = __ballot(coll);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
    const uint32_t slot = 64 * w + rank;
    if (coll) {
        // the operand bits move as integers (a float vector in between may be
        // copied with packed-float moves, which are disabled in this build)
        const uint64_t lo4 = __builtin_bit_cast(uint64_t, mlp_pack4(in[0], in[1], in[2], in[3]));
        const uint64_t hi4 = __builtin_bit_cast(uint64_t, mlp_pack4(in[4], in[5], 1.0f, 0.0f));
        X.in[slot] = u32v4{static_cast<uint32_t>(lo4), static_cast<uint32_t>(lo4 >> 32),
                           static_cast<uint32_t>(hi4), static_cast<uint32_t>(hi4 >> 32)};
    }
    uint32_t* cnt = X.count[step & 1];
    if (lane == 0) cnt[w] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    // start[v]: number of colliding ions in wavefronts before v
    uint32_t start[WAVES + 1];
    start[0] = 0;
#pragma unroll
    for (int v = 0; v < WAVES; ++v) start[v + 1] = start[v] + __builtin_amdgcn_readfirstlane(cnt[v]);
    const uint32_t total = start[WAVES];
    if (total == 0) return false;
    const uint32_t nb = (total + 63) >> 6;
    const uint32_t b = (w + nw - rot % nw) % nw;
    if (b < nb) {
        // slot of colliding ion g (g < total) in the exchange: its wavefront
        // is the number of wavefront starts past 0 at or below g
        auto slot_of = [&](uint32_t g) {
            uint32_t ww = 0, base = 0;
#pragma unroll
            for (int v = 1; v < WAVES; ++v) {
                const bool past = g >= start[v];
                ww += past ? 1u : 0u;
                base = past ? start[v] : base;
            }
            return 64 * ww + (g - base);
        };
        const uint32_t half = lane >> 5;
        const uint32_t g_own = 64 * b + lane, g_par = g_own ^ 32u;
        const bool have_own = g_own < total, have_par = g_par < total;
        const uint32_t s_own = have_own ? slot_of(g_own) : 0u;
        const uint32_t s_par = have_par ? slot_of(g_par) : 0u;
        // lane l carries operand slots 4*half .. 4*half+3: dwords 2*half, 2*half+1
        const uint32_t* xin = reinterpret_cast<const uint32_t*>(X.in);
        uint32_t o0 = 0u, o1 = 0u, q0 = 0u, q1 = 0u;
        if (have_own) {
            o0 = xin[4 * s_own + 2 * half];
            o1 = xin[4 * s_own + 2 * half + 1];
        }
        if (have_par) {
            q0 = xin[4 * s_par + 2 * half];
            q1 = xin[4 * s_par + 2 * half + 1];
        }
        const bf16x4 own = __builtin_bit_cast(bf16x4, static_cast<uint64_t>(o0) | (static_cast<uint64_t>(o1) << 32));
        const bf16x4 par = __builtin_bit_cast(bf16x4, static_cast<uint64_t>(q0) | (static_cast<uint64_t>(q1) << 32));
        // block 0 = ions 64 b + 0..31, block 1 = ions 64 b + 32..63; the
        // blocks go through the network one at a time (fewer live registers),
        // and a block past the last colliding ion is skipped
        const bf16x4 X0 = half == 0 ? own : par;
        const bf16x4 X1 = half == 0 ? par : own;
        const uint32_t n_blocks = total > 64 * b + 32 ? 2u : 1u;
        // a wave in its matrix-core section runs at raised priority so that
        // its MFMA chain keeps issuing while the other wave of the SIMD fills
        // the gaps with vector work
        __builtin_amdgcn_s_setprio(1);
        // lanes < 32: block 0 is ion g_own, block 1 is ion g_par = g_own + 32
        float r[3];
        mlp_block_core(lds, X0, r);
        if (half == 0 && have_own) {
#pragma unroll
            for (int c = 0; c < kMlpOut; ++c) X.out[c][s_own] = r[c];
        }
        if (n_blocks > 1) {
            // a compiler memory fence: the second block reloads the weight
            // fragments from LDS instead of keeping the first block's 128
            // registers of them live
            asm volatile("" ::: "memory");
            mlp_block_core(lds, X1, r);
            if (half == 0 && have_par) {
#pragma unroll
                for (int c = 0; c < kMlpOut; ++c) X.out[c][s_par] = r[c];
            }
        }
        __builtin_amdgcn_s_setprio(0);
    }
    __syncthreads();
    if (coll) {
#pragma unroll
        for (int c = 0; c < kMlpOut; ++c) out[c] = X.out[c][slot];
    }
    return true;
}

}  // namespace icarion
