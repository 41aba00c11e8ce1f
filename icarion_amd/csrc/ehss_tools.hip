// ehss_tools.hip - device kernels around EHSS geometry:
//   * synth_clusters_kernel: random rigid atom clusters (workload synthesis);
//   * oapa_kernel: orientation-averaged bounding-disk area over the 64 fixed
//     orientations of EHSSCollisionHandler::compute_effective_ccs
//     (EHSSCollisionHandler.cpp:474-498) - the rotations come from the host;
//   * features_kernel: cluster shape features fed to the surrogate;
//   * precompute_kernel: the offline EHSS precompute of
//     tools/ehss_offline_precompute.cpp:438-517 (sigma_eff per orientation by
//     hit fraction, mu = -n_z of the first-contact normal conditional on a hit).
// One thread per cluster (or per (cluster, orientation)); atoms are float4
// (x, y, z, r) in metres, arithmetic FP64 like the reference tool.
#include "device_common.hpp"
#include "step_params.hpp"

namespace icarion {

constexpr uint64_t kPrecomputeKey = 0x6a09e667f3bcc909ull;
constexpr uint64_t kClusterKey = 0xbb67ae8584caa73bull;

// Random rigid cluster by sequential attachment: atom j (radius U[r_min, r_max])
// touches a random earlier atom at 0.95 (r_p + r_j) along a random direction and
// may not come closer than 0.8 (r_k + r_j) to any other atom (up to 64 tries,
// the last candidate is kept otherwise); finally centred on the mean position.
__global__ void synth_clusters_kernel(float4* atoms, const uint32_t* offsets, const uint32_t* counts,
                                      uint32_t n_clusters, float r_min, float r_max, uint64_t seed) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_clusters) return;
    float4* a = atoms + offsets[c];
    const uint32_t n = counts[c];
    PhiloxStream rng(seed ^ kClusterKey, c, 0);
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (uint32_t j = 0; j < n; ++j) {
        const float rj = r_min + (r_max - r_min) * rng.uniform_f32();
        float4 cand = make_float4(0.f, 0.f, 0.f, rj);
        if (j > 0) {
            for (int attempt = 0; attempt < 64; ++attempt) {
                uint32_t p = static_cast<uint32_t>(rng.uniform_f32() * static_cast<float>(j));
                p = p >= j ? j - 1 : p;
                const float4 ap = a[p];
                const float u = 2.f * rng.uniform_f32() - 1.f;
                const float ph = rng.uniform_f32();
                const float st = sqrtf(fmaxf(0.f, 1.f - u * u));
                const float d = 0.95f * (ap.w + rj);
                cand = make_float4(ap.x + d * st * __builtin_amdgcn_cosf(ph),
                                   ap.y + d * st * __builtin_amdgcn_sinf(ph), ap.z + d * u, rj);
                bool ok = true;
                for (uint32_t k = 0; k < j && ok; ++k) {
                    if (k == p) continue;
                    const float4 ak = a[k];
                    const float dx = cand.x - ak.x, dy = cand.y - ak.y, dz = cand.z - ak.z;
                    const float lim = 0.8f * (ak.w + rj);
                    ok = dx * dx + dy * dy + dz * dz >= lim * lim;
                }
                if (ok) break;
            }
        }
        a[j] = cand;
        sx += cand.x; sy += cand.y; sz += cand.z;
    }
    const float inv = 1.f / static_cast<float>(n);
    sx *= inv; sy *= inv; sz *= inv;
    for (uint32_t j = 0; j < n; ++j) {
        float4 v = a[j];
        v.x -= sx; v.y -= sy; v.z -= sz;
        a[j] = v;
    }
}

// sigma_oapa = mean over the 64 orientations R_k of pi * (max_j (|R_k c_j|_xy + r_j) + r_gas)^2
__global__ void oapa_kernel(const float4* atoms, const uint32_t* offsets, const uint32_t* counts,
                            uint32_t n_clusters, const double* rot64, double r_gas,
                            double* sigma_out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_clusters) return;
    const float4* a = atoms + offsets[c];
    const uint32_t n = counts[c];
    double A_sum = 0.0;
    for (int k = 0; k < 64; ++k) {
        const double* R = rot64 + 9 * k;
        double Rmax = 0.0;
        for (uint32_t j = 0; j < n; ++j) {
            const float4 v = a[j];
            const double cx = v.x, cy = v.y, cz = v.z;
            const double x = R[0] * cx + R[1] * cy + R[2] * cz;
            const double y = R[3] * cx + R[4] * cy + R[5] * cz;
            const double r_eff = sqrt(x * x + y * y) + static_cast<double>(v.w) + r_gas;
            Rmax = fmax(Rmax, r_eff);
        }
        A_sum += kPi * Rmax * Rmax;
    }
    sigma_out[c] = A_sum / 64.0;
}

__global__ void features_kernel(const float4* atoms, const uint32_t* offsets, const uint32_t* counts,
                                uint32_t n_clusters, double r_gas, const double* sigma_oapa,
                                float* f) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_clusters) return;
    const float4* a = atoms + offsets[c];
    const uint32_t n = counts[c];
    double mx = 0, my = 0, mz = 0, mr = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const float4 v = a[j];
        mx += v.x; my += v.y; mz += v.z; mr += v.w;
    }
    mx /= n; my /= n; mz /= n; mr /= n;
    double rg2 = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const float4 v = a[j];
        const double dx = v.x - mx, dy = v.y - my, dz = v.z - mz;
        rg2 += dx * dx + dy * dy + dz * dz;
    }
    const double b_eff = sqrt(sigma_oapa[c] / kPi);
    f[3 * c] = static_cast<float>(log(static_cast<double>(n)) / log(500.0));
    f[3 * c + 1] = static_cast<float>(sqrt(rg2 / n) / b_eff);
    f[3 * c + 2] = static_cast<float>((mr + r_gas) / b_eff);
}

// first contact along +z of a ray through (ox, oy) (tools/ehss_offline_precompute.cpp:162-213);
// atoms rotated on the fly by R (same products as storing them rotated)
__device__ bool precompute_contact(const float4* a, uint32_t n, const double R[9], double ox, double oy,
                                   double r_gas, double& mu) {
    bool hit = false;
    double best_s = 0.0, nz_best = 0.0;
    for (uint32_t j = 0; j < n; ++j) {
        const float4 v = a[j];
        const double cx = v.x, cy = v.y, cz = v.z;
        const double rax = R[0] * cx + R[1] * cy + R[2] * cz;
        const double ray = R[3] * cx + R[4] * cy + R[5] * cz;
        const double raz = R[6] * cx + R[7] * cy + R[8] * cz;
        const double Rsum = static_cast<double>(v.w) + r_gas;
        const double Rsum2 = Rsum * Rsum;
        const double rx = ox - rax, ry = oy - ray, rz = 0.0 - raz;
        const double dmin2 = rx * rx + ry * ry;
        if (dmin2 <= Rsum2) {
            const double h = sqrt(fmax(0.0, Rsum2 - dmin2));
            const double s_hit = rz - h;
            const double pz = rz - s_hit;
            const double phr2 = rx * rx + ry * ry + pz * pz;
            double nz;
            if (phr2 <= kMinContactDistSq) {
                nz = -1.0;
            } else {
                nz = pz * (1.0 / sqrt(phr2));
                if (nz > 0.0) nz = -nz;
            }
            if (!hit || s_hit < best_s) {
                hit = true;
                best_s = s_hit;
                nz_best = nz;
            }
        }
    }
    if (hit) mu = fmin(1.0, fmax(-1.0, -nz_best));
    return hit;
}

// One thread per (cluster, orientation).
__global__ void precompute_kernel(const float4* atoms, const uint32_t* offsets, const uint32_t* counts,
                                  uint32_t cluster_begin, uint32_t n_clusters, uint32_t n_orient,
                                  uint32_t n_sigma, uint32_t n_mu, uint32_t max_tries, uint64_t seed,
                                  double r_gas, double* sigma_out, double* mu_out) {
    const uint64_t tid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= static_cast<uint64_t>(n_clusters) * n_orient) return;
    const uint32_t cl = static_cast<uint32_t>(tid / n_orient);
    const uint32_t k = static_cast<uint32_t>(tid % n_orient);
    const uint32_t c = cluster_begin + cl;
    PhiloxStream rng(seed ^ kPrecomputeKey, c, k);
    // sample_random_quaternion + quaternion_to_rotation (normalised)
    const double u1 = rng.uniform_f64();
    const double u2 = rng.uniform_f64();
    const double u3 = rng.uniform_f64();
    double x = sqrt(1.0 - u1) * sin(2.0 * kPi * u2);
    double y = sqrt(1.0 - u1) * cos(2.0 * kPi * u2);
    double z = sqrt(u1) * sin(2.0 * kPi * u3);
    double w = sqrt(u1) * cos(2.0 * kPi * u3);
    const double nrm = sqrt(w * w + x * x + y * y + z * z);
    w /= nrm; x /= nrm; y /= nrm; z /= nrm;
    double R[9];
    R[0] = 1.0 - 2.0 * (y * y + z * z); R[1] = 2.0 * (x * y - z * w); R[2] = 2.0 * (x * z + y * w);
    R[3] = 2.0 * (x * y + z * w); R[4] = 1.0 - 2.0 * (x * x + z * z); R[5] = 2.0 * (y * z - x * w);
    R[6] = 2.0 * (x * z - y * w); R[7] = 2.0 * (y * z + x * w); R[8] = 1.0 - 2.0 * (x * x + y * y);
    const float4* a = atoms + offsets[c];
    const uint32_t n = counts[c];
    double b_base = 0.0;
    for (uint32_t j = 0; j < n; ++j) {
        const float4 v = a[j];
        const double cx = v.x, cy = v.y, cz = v.z;
        const double rax = R[0] * cx + R[1] * cy + R[2] * cz;
        const double ray = R[3] * cx + R[4] * cy + R[5] * cz;
        b_base = fmax(b_base, sqrt(rax * rax + ray * ray) + static_cast<double>(v.w));
    }
    const double b_max = b_base + r_gas;
    const double area = kPi * b_max * b_max;
    uint32_t hits = 0;
    double mu = 0.0;
    for (uint32_t s = 0; s < n_sigma; ++s) {
        const double v1 = rng.uniform_f64();
        const double v2 = rng.uniform_f64();
        const double r = b_max * sqrt(v1);
        const double phi = 2.0 * kPi * v2;
        if (precompute_contact(a, n, R, r * cos(phi), r * sin(phi), r_gas, mu)) ++hits;
    }
    sigma_out[tid] = n_sigma ? (static_cast<double>(hits) / static_cast<double>(n_sigma)) * area : 0.0;
    double* row = mu_out + tid * n_mu;
    uint32_t got = 0, tries = 0;
    while (got < n_mu && tries < max_tries) {
        ++tries;
        const double v1 = rng.uniform_f64();
        const double v2 = rng.uniform_f64();
        const double r = b_max * sqrt(v1);
        const double phi = 2.0 * kPi * v2;
        if (precompute_contact(a, n, R, r * cos(phi), r * sin(phi), r_gas, mu)) row[got++] = mu;
    }
    for (; got < n_mu; ++got) row[got] = __builtin_nan("");
}

hipError_t launch_synth_clusters(float4* atoms, const uint32_t* offsets, const uint32_t* counts,
                                 uint32_t n, float r_min, float r_max, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL(synth_clusters_kernel, dim3((n + 127) / 128), dim3(128), 0, s, atoms, offsets,
                       counts, n, r_min, r_max, seed);
    return hipGetLastError();
}

hipError_t launch_oapa(const float4* atoms, const uint32_t* offsets, const uint32_t* counts, uint32_t n,
                       const double* rot64, double r_gas, double* sigma, hipStream_t s) {
    hipLaunchKernelGGL(oapa_kernel, dim3((n + 127) / 128), dim3(128), 0, s, atoms, offsets, counts, n,
                       rot64, r_gas, sigma);
    return hipGetLastError();
}

hipError_t launch_features(const float4* atoms, const uint32_t* offsets, const uint32_t* counts,
                           uint32_t n, double r_gas, const double* sigma, float* f, hipStream_t s) {
    hipLaunchKernelGGL(features_kernel, dim3((n + 127) / 128), dim3(128), 0, s, atoms, offsets, counts,
                       n, r_gas, sigma, f);
    return hipGetLastError();
}

hipError_t launch_precompute(const float4* atoms, const uint32_t* offsets, const uint32_t* counts,
                             uint32_t cluster_begin, uint32_t n_clusters, uint32_t n_orient,
                             uint32_t n_sigma, uint32_t n_mu, uint32_t max_tries, uint64_t seed,
                             double r_gas, double* sigma, double* mu, hipStream_t s) {
    const uint64_t nt = static_cast<uint64_t>(n_clusters) * n_orient;
    hipLaunchKernelGGL(precompute_kernel, dim3(static_cast<uint32_t>((nt + 127) / 128)), dim3(128), 0, s,
                       atoms, offsets, counts, cluster_begin, n_clusters, n_orient, n_sigma, n_mu,
                       max_tries, seed, r_gas, sigma, mu);
    return hipGetLastError();
}

}  // namespace icarion
