// device_common.hpp - device-side building blocks of the gfx950 ion-step
// kernels: 3-vectors, the Philox4x32-10 counter RNG and per-(ion, step)
// random streams.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icarion {

// ---- physical constants (reference src/utils/constants.h:13-65) ----------
constexpr double kPi = 3.14159265358979323846;
constexpr double kBoltzmann = 1.380649e-23;
constexpr double kLoschmidt = 2.6867811e25;
constexpr double kEpsilon0 = 8.854187817e-12;
constexpr double kStpTemp = 273.15;
// thresholds (HSSCollisionHandler.cpp:20, CollisionKernels.h:172-178,
// CylindricalGeometry.h:245, AnalyticalFieldModel.cpp:12)
constexpr double kMinRelativeVelocity = 1e-10;
constexpr double kMinVelocityMag = 1e-12;
constexpr double kMinContactDistSq = 1e-24;
constexpr int kEhssMaxAttempts = 256;
constexpr int kBmaxExpansionAttempt = 64;
constexpr double kBmaxExpansionFactor = 1.5;
constexpr double kGeomEpsilon = 1e-12;
constexpr double kMinVoltage = 1e-12;

template <typename T>
struct vec3 {
    T x, y, z;
};
template <typename T>
__device__ __forceinline__ vec3<T> operator+(vec3<T> a, vec3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T>
__device__ __forceinline__ vec3<T> operator-(vec3<T> a, vec3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T>
__device__ __forceinline__ vec3<T> operator*(vec3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T>
__device__ __forceinline__ T dot(vec3<T> a, vec3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T>
__device__ __forceinline__ vec3<T> cross(vec3<T> a, vec3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ vec3<double> mat_mul(const double m[9], vec3<double> v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z,
            m[3] * v.x + m[4] * v.y + m[5] * v.z,
            m[6] * v.x + m[7] * v.y + m[8] * v.z};
}

// ---- Philox4x32-10 (Salmon et al., SC'11), published round constants -----
struct u32x4 {
    uint32_t v[4];
};

__device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    u32x4 o;
    o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

// 53-bit uniform on (0, 1] from two words: k*2^-53 + 2^-54, k = hi[31:5]:lo[31:6].
__device__ __forceinline__ double uniform53(uint32_t hi, uint32_t lo) {
    const uint64_t k = (static_cast<uint64_t>(hi >> 5) << 26) | static_cast<uint64_t>(lo >> 6);
    return static_cast<double>(k) * 0x1.0p-53 + 0x1.0p-54;
}

// 24-bit uniform on (0, 1) from one word.
__device__ __forceinline__ float uniform24(uint32_t w) {
    return static_cast<float>(w >> 8) * 0x1.0p-24f + 0x1.0p-25f;
}

// Per-(ion, step) stream, counter = (block, ion, step_lo, step_hi), key =
// (seed_lo, seed_hi).  FP64 draws take two words (the oracle's PhiloxStream
// hands out the identical sequence); FP32 draws take one word.
struct PhiloxStream {
    uint32_t k0, k1, ion, s_lo, s_hi;
    uint32_t block;
    uint32_t buf[4];
    int pos;
    bool n_have;       // normal_f64: the second value of the last pair is pending
    double n_cached;

    __device__ __forceinline__ PhiloxStream(uint64_t seed, uint32_t ion_index, uint64_t step)
        : k0(static_cast<uint32_t>(seed)), k1(static_cast<uint32_t>(seed >> 32)), ion(ion_index),
          s_lo(static_cast<uint32_t>(step)), s_hi(static_cast<uint32_t>(step >> 32)), block(0),
          pos(4), n_have(false), n_cached(0.0) {}

    __device__ __forceinline__ void refill() {
        const u32x4 o = philox4x32_10(block, ion, s_lo, s_hi, k0, k1);
        buf[0] = o.v[0]; buf[1] = o.v[1]; buf[2] = o.v[2]; buf[3] = o.v[3];
        ++block;
        pos = 0;
    }
    // The words are selected rather than indexed by pos, so that the block
    // stays in registers where the compiler cannot resolve pos (draws inside
    // data-dependent loops); an indexed buffer goes to private memory there.
    // A stream takes FP64 draws only (pos 0 or 2) or FP32 draws only.
    __device__ __forceinline__ double uniform_f64() {
        if (pos >= 4) refill();
        const bool second = pos >= 2;
        const double u = uniform53(second ? buf[2] : buf[0], second ? buf[3] : buf[1]);
        pos += 2;
        return u;
    }
    // Standard normal by Box-Muller on two FP64 draws, the pair's second value
    // kept for the next call (the oracle's stream and std::normal_distribution
    // hand out both values of a pair), so consecutive collision substeps share
    // pairs as the engine's one stream per ion does.
    __device__ __forceinline__ double normal_f64() {
        if (n_have) {
            n_have = false;
            return n_cached;
        }
        const double u1 = uniform_f64();
        const double u2 = uniform_f64();
        const double r = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincos(2.0 * 3.14159265358979323846 * u2, &sn, &cs);
        n_cached = r * sn;
        n_have = true;
        return r * cs;
    }
    __device__ __forceinline__ float uniform_f32() {
        if (pos >= 4) refill();
        const uint32_t w = pos == 0 ? buf[0] : pos == 1 ? buf[1] : pos == 2 ? buf[2] : buf[3];
        const float u = uniform24(w);
        pos += 1;
        return u;
    }
};

}  // namespace icarion
