// metrics_kernels.hip - arrival-time statistics and histogram over the
// resident ions (the in-kernel first crossing of the detector plane).
//
// The reference's analysis takes the arrival times, builds histogram edges
// linspace(min, max, bins + 1) and bins them with numpy
// (analysis/physics.py:96-104, 119-145).  Here:
//   arrival_partials_kernel: per workgroup, count / sum / sum of squares /
//     min / max of the finite arrival times (wave shuffles, then LDS), one
//     record per workgroup; the host sums the records in workgroup order, so
//     the result does not depend on scheduling;
//   arrival_histogram_kernel: equal-width bins with numpy's own bin rule for
//     them (index = floor(((x - lo) / (hi - lo)) * n), the top edge into the
//     last bin, then one step down or up where x falls outside the edges
//     i * ((hi - lo) / n) + lo, the top one = hi), counted in LDS and added
//     to the global bins once per workgroup.
#include <hip/hip_runtime.h>
#include <stdint.h>

// the bin edges must round as numpy's linspace does (i * step, then + lo)
#pragma clang fp contract(off)

namespace icarion {

struct ArrivalPartial {
    double sum, sumsq, min, max;
    unsigned long long count, pad;
};

constexpr int kMetricThreads = 256;
constexpr uint32_t kMaxHistBins = 8192;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_min_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__global__ void __launch_bounds__(kMetricThreads) arrival_partials_kernel(const double* __restrict__ arrival, uint32_t n,
                                                                          ArrivalPartial* __restrict__ out) {
    __shared__ double s_sum[kMetricThreads / 64], s_sq[kMetricThreads / 64], s_min[kMetricThreads / 64],
        s_max[kMetricThreads / 64];
    __shared__ unsigned long long s_cnt[kMetricThreads / 64];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double x = (i < n) ? arrival[i] : NAN;
    const bool ok = isfinite(x);
    const double v = ok ? x : 0.0;
    double sum = wave_sum_d(v);
    double sq = wave_sum_d(v * v);
    double mn = wave_min_d(ok ? x : INFINITY);
    double mx = wave_max_d(ok ? x : -INFINITY);
    const unsigned long long cnt = __popcll(__ballot(ok));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_sum[wave] = sum;
        s_sq[wave] = sq;
        s_min[wave] = mn;
        s_max[wave] = mx;
        s_cnt[wave] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ArrivalPartial p{0.0, 0.0, INFINITY, -INFINITY, 0ull, 0ull};
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
            p.sum += s_sum[w];
            p.sumsq += s_sq[w];
            p.min = fmin(p.min, s_min[w]);
            p.max = fmax(p.max, s_max[w]);
            p.count += s_cnt[w];
        }
        out[blockIdx.x] = p;
    }
}

__global__ void __launch_bounds__(kMetricThreads) arrival_histogram_kernel(const double* __restrict__ arrival, uint32_t n,
                                                                           double lo, double hi, uint32_t n_bins,
                                                                           unsigned long long* __restrict__ bins) {
    __shared__ unsigned int s_bins[kMaxHistBins];
    for (uint32_t b = threadIdx.x; b < n_bins; b += blockDim.x) s_bins[b] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double x = arrival[i];
        if (isfinite(x) && x >= lo && x <= hi) {
            const double step = (hi - lo) / static_cast<double>(n_bins);
            long long k = static_cast<long long>(((x - lo) / (hi - lo)) * static_cast<double>(n_bins));
            if (k >= static_cast<long long>(n_bins)) k = n_bins - 1;
            if (k < 0) k = 0;
            auto edge = [&](long long e) { return e >= static_cast<long long>(n_bins) ? hi : e * step + lo; };
            if (x < edge(k)) --k;
            else if (k != static_cast<long long>(n_bins) - 1 && x >= edge(k + 1)) ++k;
            if (k >= 0 && k < static_cast<long long>(n_bins)) atomicAdd(&s_bins[k], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < n_bins; b += blockDim.x)
        if (s_bins[b]) atomicAdd(&bins[b], static_cast<unsigned long long>(s_bins[b]));
}

__global__ void zero_bins_kernel(unsigned long long* bins, uint32_t n_bins) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n_bins; b += gridDim.x * blockDim.x) bins[b] = 0ull;
}

uint32_t arrival_blocks(uint32_t n) { return (n + kMetricThreads - 1) / kMetricThreads; }

hipError_t launch_arrival_partials(const double* arrival, uint32_t n, void* partials, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(arrival_partials_kernel, dim3(arrival_blocks(n)), dim3(kMetricThreads), 0, s, arrival, n,
                       static_cast<ArrivalPartial*>(partials));
    return hipGetLastError();
}

hipError_t launch_arrival_histogram(const double* arrival, uint32_t n, double lo, double hi, uint32_t n_bins,
                                    unsigned long long* bins, hipStream_t s) {
    if (n_bins == 0 || n_bins > kMaxHistBins) return hipErrorInvalidValue;
    hipLaunchKernelGGL(zero_bins_kernel, dim3((n_bins + 255) / 256), dim3(256), 0, s, bins, n_bins);
    if (n > 0)
        hipLaunchKernelGGL(arrival_histogram_kernel, dim3(arrival_blocks(n)), dim3(kMetricThreads), 0, s, arrival, n,
                           lo, hi, n_bins, bins);
    return hipGetLastError();
}

}  // namespace icarion
