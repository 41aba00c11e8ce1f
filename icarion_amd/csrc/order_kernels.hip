// order_kernels.hip - the ion order of a field-array launch: ions sorted by the
// 4 x 4 x 4-node brick of the grid that holds them, so that the lanes of a
// workgroup interpolate in neighbouring cells and their corner gathers share
// cache lines (an unsorted cloud scatters every wavefront's gathers over the
// whole grid).  Dead ions sort last.  The order only changes which lane runs an
// ion: every ion keeps its index, so the results do not depend on it.
//
// Also the epoch shuffle of the surrogate training: a random permutation as
// the sort of one Philox key per sample.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include <stdint.h>

#include "device_common.hpp"

namespace icarion {

namespace {

__global__ void __launch_bounds__(256) brick_keys_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                         const double* __restrict__ z,
                                                         const uint8_t* __restrict__ flags, uint32_t n, double x0,
                                                         double y0, double z0, double sx, double sy, double sz,
                                                         uint32_t nbx, uint32_t nby, uint32_t nbz,
                                                         uint32_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    idx[i] = i;
    // born but inactive: dead; unborn ions keep their place by position
    const uint8_t f = flags[i];
    if ((f & 2u) && !(f & 1u)) {
        keys[i] = nbx * nby * nbz;
        return;
    }
    auto brick = [](double v, double v0, double s, uint32_t nb) -> uint32_t {
        const double b = (v - v0) * s;   // in bricks
        if (!(b > 0.0)) return 0u;       // also NaN
        const uint32_t k = b >= static_cast<double>(nb) ? nb - 1u : static_cast<uint32_t>(b);
        return k;
    };
    keys[i] = brick(x[i], x0, sx, nbx) + nbx * (brick(y[i], y0, sy, nby) + nby * brick(z[i], z0, sz, nbz));
}

// keys[i] = Philox4x32-10 of counter (i, epoch, 0, 0) under the key (seed), idx[i] = i
__global__ void __launch_bounds__(256) shuffle_keys_kernel(uint32_t n, uint64_t seed, uint32_t epoch,
                                                           uint32_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = philox4x32_10(i, epoch, 0u, 0u, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)).v[0];
    idx[i] = i;
}

}  // namespace

// perm[0..n) <- a random permutation of 0..n-1 that depends only on (seed,
// epoch): the indices ordered by their Philox keys (ties by index, the sort is
// stable).  keys_a, keys_b, idx: n entries each; temp: brick_sort_temp_bytes(n).
hipError_t random_permutation(uint32_t n, uint64_t seed, uint32_t epoch, uint32_t* keys_a, uint32_t* keys_b,
                              uint32_t* idx, uint32_t* perm, void* temp, size_t temp_bytes, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(shuffle_keys_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, seed, epoch, keys_a, idx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_a, keys_b, idx, perm, static_cast<int>(n), 0, 32,
                                              s);
}

// Scratch bytes of sort_ions_by_brick for n ions (hipCUB's temporary storage).
size_t brick_sort_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, static_cast<const uint32_t*>(nullptr),
                                       static_cast<uint32_t*>(nullptr), static_cast<const uint32_t*>(nullptr),
                                       static_cast<uint32_t*>(nullptr), static_cast<int>(n));
    return bytes;
}

// order[0..n) <- the ion indices sorted by brick key.  extent: the grid's first
// node and the reciprocal brick size per axis; nb: bricks per axis.  keys_a,
// keys_b, idx: n entries each; temp: brick_sort_temp_bytes(n) bytes.
hipError_t sort_ions_by_brick(const double* x, const double* y, const double* z, const uint8_t* flags, uint32_t n,
                              const double origin[3], const double inv_brick[3], const uint32_t nb[3],
                              uint32_t* keys_a, uint32_t* keys_b, uint32_t* idx, uint32_t* order, void* temp,
                              size_t temp_bytes, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(brick_keys_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, y, z, flags, n, origin[0],
                       origin[1], origin[2], inv_brick[0], inv_brick[1], inv_brick[2], nb[0], nb[1], nb[2], keys_a,
                       idx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) <= static_cast<unsigned long long>(nb[0]) * nb[1] * nb[2]) ++end_bit;
    return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_a, keys_b, idx, order, static_cast<int>(n), 0,
                                              static_cast<int>(end_bit), s);
}

}  // namespace icarion
