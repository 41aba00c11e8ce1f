// icarion_oracle.cpp - CPU FP64 restatement of the reference batch ion-step.
//
// TEST INFRASTRUCTURE ONLY: the parity checker and the host-CPU baseline.
// It is never linked into, loaded by, or called from the product library.
//
// Each function below restates the reference function cited next to it,
// keeping its operation order so that, driven by the same random numbers,
// results agree to rounding.  Two random sources:
//   * Philox4x32-10 streams keyed by (rng_seed, ion) with the step number in
//     the counter - the streams the device kernels draw, so GPU and oracle
//     trajectories can be compared ion by ion;
//   * std::mt19937_64 + uniform_real_distribution / normal_distribution
//     seeded rng_seed + i - the reference's PhysicsRng (CollisionTypes.cpp:8-20,
//     RngUtils.cpp:51-53), so the oracle can be pinned bit for bit against the
//     reference's own compiled kernels.
//
// Parity status of each part is listed in the file header of
// tests/test_oracle_pins.py (what pins it and how).
#include "icarion_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---- constants (src/utils/constants.h:13-65) ------------------------------
constexpr double kPi = 3.14159265358979323846;
constexpr double kBoltzmann = 1.380649e-23;
constexpr double kLoschmidt = 2.6867811e25;
constexpr double kEpsilon0 = 8.854187817e-12;
constexpr double kStpTemp = 273.15;
// HSSCollisionHandler.cpp:20, EHSSCollisionHandler.cpp:20
constexpr double kMinRelativeVelocity = 1e-10;
// CollisionKernels.h:172-178
constexpr double kMinVelocityMag = 1e-12;
constexpr double kMinContactDistSq = 1e-24;
constexpr int kEhssMaxAttempts = 256;
constexpr int kBmaxExpansionAttempt = 64;
constexpr double kBmaxExpansionFactor = 1.5;
// CylindricalGeometry.h:245, OrbitrapGeometry.h:231
constexpr double kGeomEpsilon = 1e-12;
// AnalyticalFieldModel.cpp:12
constexpr double kMinVoltage = 1e-12;

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {  // mathUtils.h:32-35
    double n = norm(a);
    return n > 0.0 ? a / n : V3{0, 0, 0};
}
inline V3 mat_mul(const double m[9], V3 v) {  // mathUtils.h:88-94 (row-major)
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z,
            m[3] * v.x + m[4] * v.y + m[5] * v.z,
            m[6] * v.x + m[7] * v.y + m[8] * v.z};
}
inline bool finite3(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random
// numbers: as easy as 1, 2, 3"); round multipliers and Weyl key increments
// are the published constants of the algorithm.
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

inline void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(kPhiloxM0) * c0;
        const uint64_t p1 = static_cast<uint64_t>(kPhiloxM1) * c2;
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n1 = lo1;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        const uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 53-bit uniform on (0, 1]: k * 2^-53 + 2^-54 with k from the top 27 bits of
// hi and the top 26 bits of lo.  Never 0, so log(u) is always finite.
inline double uniform_from_words(uint32_t hi, uint32_t lo) {
    const uint64_t k = (static_cast<uint64_t>(hi >> 5) << 26) | static_cast<uint64_t>(lo >> 6);
    return static_cast<double>(k) * 0x1.0p-53 + 0x1.0p-54;
}

// Per-(ion, step) Philox stream: counter = (block, ion, step_lo, step_hi),
// key = (seed_lo, seed_hi).  uniform01() consumes two words; normal() is
// Box-Muller and caches the sine branch.
class PhiloxStream {
public:
    PhiloxStream(uint64_t seed, uint32_t ion, uint64_t step) {
        key_[0] = static_cast<uint32_t>(seed);
        key_[1] = static_cast<uint32_t>(seed >> 32);
        ctr_[0] = 0;
        ctr_[1] = ion;
        ctr_[2] = static_cast<uint32_t>(step);
        ctr_[3] = static_cast<uint32_t>(step >> 32);
        pos_ = 4;
    }
    double uniform01() {
        if (pos_ >= 4) {
            philox(ctr_, key_, buf_);
            ctr_[0] += 1;
            pos_ = 0;
        }
        const double u = uniform_from_words(buf_[pos_], buf_[pos_ + 1]);
        pos_ += 2;
        return u;
    }
    double normal() {
        if (has_cached_) {
            has_cached_ = false;
            return cached_;
        }
        const double u1 = uniform01();
        const double u2 = uniform01();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double a = 2.0 * kPi * u2;
        cached_ = r * std::sin(a);
        has_cached_ = true;
        return r * std::cos(a);
    }

private:
    uint32_t key_[2];
    uint32_t ctr_[4];
    uint32_t buf_[4];
    int pos_;
    bool has_cached_ = false;
    double cached_ = 0.0;
};

// PhysicsRng (CollisionTypes.cpp:8-20): mt19937_64 + uniform_real_distribution
// + normal_distribution.
class MtRng {
public:
    explicit MtRng(uint64_t seed) : eng_(seed), uni_(0.0, 1.0), norm_(0.0, 1.0) {}
    double uniform01() { return uni_(eng_); }
    double normal() { return norm_(eng_); }

private:
    std::mt19937_64 eng_;
    std::uniform_real_distribution<double> uni_;
    std::normal_distribution<double> norm_;
};

// ---- fields -------------------------------------------------------------

// The waveform types of WaveformConfig.h:33-200 and ArbitraryWaveform::evaluate
// (WaveformConfig.cpp:10-91).
double waveform_value(const icarion_waveform_desc& w, double t) {
    const double* p = w.p;
    switch (w.type) {
        case ICARION_WAVEFORM_CONSTANT: return p[0];
        case ICARION_WAVEFORM_LINEAR: {
            const double start_value = p[0], end_value = p[1], start_time = p[2], end_time = p[3];
            if (t < start_time) return start_value;
            if (t >= end_time) return w.clamp ? end_value : start_value;
            const double duration = end_time - start_time;
            const double t_norm = (t - start_time) / duration;
            return start_value + (end_value - start_value) * t_norm;
        }
        case ICARION_WAVEFORM_QUADRATIC: {
            const double a = p[0], b = p[1], c = p[2];
            if (t < p[3] || t > p[4]) return a;
            return a + b * t + c * t * t;
        }
        case ICARION_WAVEFORM_EXPONENTIAL: {
            const double offset = p[0], amplitude = p[1], rate = p[2], start = p[3], end = p[4];
            auto value_at = [&](double rel_t) { return offset + amplitude * std::exp(rate * rel_t); };
            if (t < start) return value_at(0.0);
            if (t >= end) return w.clamp ? value_at(end - start) : value_at(0.0);
            return value_at(t - start);
        }
        case ICARION_WAVEFORM_SINUSOIDAL: {
            const double omega = 2.0 * kPi * p[2];
            return p[0] + p[1] * std::sin(omega * t + p[3]);
        }
        case ICARION_WAVEFORM_PWM: {
            const double low = p[0], high = p[1], f = p[2], duty = p[3], ph = p[4];
            if (f <= 0.0) return low;
            auto eval_at = [&](double t_rel) {
                const double period = 1.0 / f;
                const double phase_shift = ph / (2.0 * kPi * f);
                double phase = std::fmod(t_rel + phase_shift, period);
                if (phase < 0.0) phase += period;
                return (phase < duty * period) ? high : low;
            };
            if (t < p[5]) return low;
            if (t >= p[6]) return w.clamp ? eval_at(p[6] - p[5]) : low;
            return eval_at(t - p[5]);
        }
        case ICARION_WAVEFORM_PULSED:
            return (t >= p[2] && t < p[2] + p[3]) ? p[1] : p[0];
        default: {
            const double* ts = w.times_s;
            const double* vs = w.values;
            const size_t n = w.n_points;
            if (t <= ts[0]) return vs[0];
            if (t >= ts[n - 1]) return vs[n - 1];
            const size_t idx = static_cast<size_t>(std::lower_bound(ts, ts + n, t) - ts);
            if (idx == 0) return vs[0];
            const size_t i0 = idx - 1, i1 = idx;
            const double t0 = ts[i0], t1 = ts[i1], v0 = vs[i0], v1 = vs[i1];
            if (w.interp == ICARION_INTERP_STEP) return v0;
            if (w.interp == ICARION_INTERP_LINEAR) {
                const double alpha = (t - t0) / (t1 - t0);
                return v0 + alpha * (v1 - v0);
            }
            const double alpha = (t - t0) / (t1 - t0);
            const double m0 = (i0 == 0) ? (v1 - v0) / (t1 - t0) : (v1 - vs[i0 - 1]) / (t1 - ts[i0 - 1]);
            const double m1 = (i1 == n - 1) ? (v1 - v0) / (t1 - t0) : (vs[i1 + 1] - v0) / (ts[i1 + 1] - t0);
            // Hermite basis in the reference's operation order
            const double h00 = 2 * alpha * alpha * alpha - 3 * alpha * alpha + 1;
            const double h10 = alpha * alpha * alpha - 2 * alpha * alpha + alpha;
            const double h01 = -2 * alpha * alpha * alpha + 3 * alpha * alpha;
            const double h11 = alpha * alpha * alpha - alpha * alpha;
            const double dt = t1 - t0;
            return h00 * v0 + h10 * dt * m0 + h01 * v1 + h11 * dt * m1;
        }
    }
}

// eval_value of a voltage slot: its waveform when one is set, else the constant.
double slot_value(const oracle_field* f, int slot, double constant, double t) {
    return (f && f->waveforms[slot]) ? waveform_value(*f->waveforms[slot], t) : constant;
}

// AnalyticalFieldModel::compute_ims_field (AnalyticalFieldModel.cpp:136-161).
// Absent values evaluate to 0 (the has_value() guards).
V3 ims_field(const icarion_domain_desc& d, V3 pos, double t, const oracle_field* f) {
    V3 E{0.0, 0.0, 0.0};
    const double dc_axial = slot_value(f, ICARION_SLOT_DC_AXIAL, d.dc_axial_V, t);
    const double dc_quad = slot_value(f, ICARION_SLOT_DC_QUAD, d.dc_quad_V, t);
    const double rf_voltage = slot_value(f, ICARION_SLOT_RF_VOLTAGE, d.rf_voltage_V, t);
    const double rf_freq = slot_value(f, ICARION_SLOT_RF_FREQUENCY, d.rf_frequency_Hz, t);
    E.z = dc_axial / d.length_m;
    if (std::fabs(rf_voltage) > kMinVoltage) {
        const double r0_sq = d.radius_m * d.radius_m;
        const double omega = 2.0 * kPi * rf_freq;
        const double U_eff = dc_quad + rf_voltage * std::cos(omega * t);
        E.x = -2.0 * pos.x * U_eff / r0_sq;
        E.y = 2.0 * pos.y * U_eff / r0_sq;
    }
    return E;
}

// AnalyticalFieldModel::compute_quadrupole_rf_field (AnalyticalFieldModel.cpp:219-248).
// Global coordinates; rf.phase_rad is not used by the reference.
V3 quadrupole_field(const icarion_domain_desc& d, V3 pos, double t, const oracle_field* f) {
    V3 E{0.0, 0.0, 0.0};
    const double rf_voltage = slot_value(f, ICARION_SLOT_RF_VOLTAGE, d.rf_voltage_V, t);
    const double rf_freq = slot_value(f, ICARION_SLOT_RF_FREQUENCY, d.rf_frequency_Hz, t);
    const double dc_quad = slot_value(f, ICARION_SLOT_DC_QUAD, d.dc_quad_V, t);
    const double dc_axial = slot_value(f, ICARION_SLOT_DC_AXIAL, d.dc_axial_V, t);
    if (std::fabs(rf_voltage) > kMinVoltage) {
        const double r0_sq = d.radius_m * d.radius_m;
        const double omega = 2.0 * kPi * rf_freq;
        const double U_eff = dc_quad + rf_voltage * std::cos(omega * t);
        E.x = -2.0 * pos.x * U_eff / r0_sq;
        E.y = 2.0 * pos.y * U_eff / r0_sq;
    }
    if (std::fabs(dc_axial) > kMinVoltage) {
        E.z = dc_axial / d.length_m;
    }
    return E;
}

// AnalyticalFieldModel::compute_orbitrap_field (AnalyticalFieldModel.cpp:188-217):
// the DC radial voltage is evaluated at t = 0 (:198), whatever the step time.
V3 orbitrap_field(const icarion_domain_desc& d, V3 gpos, const oracle_field* f) {
    const V3 p = mat_mul(d.rot_g2l, gpos - V3{d.origin_m[0], d.origin_m[1], d.origin_m[2]});
    const double r_sq = std::max(1e-18, p.x * p.x + p.y * p.y);
    const double r_char = d.radius_char_m;
    const double r_char_sq = r_char * r_char;
    const double C = 1.0 - r_char_sq / r_sq;
    const double voltage = slot_value(f, ICARION_SLOT_DC_RADIAL, d.dc_radial_V, 0.0);
    const double r_in = d.radius_in_m;
    const double r_out = d.radius_out_m;
    const double k = 2.0 * voltage / (r_char_sq * std::log(r_out / r_in) -
                                      0.5 * (r_out * r_out - r_in * r_in));
    const V3 E_local{0.5 * k * p.x * C, 0.5 * k * p.y * C, -k * p.z};
    return mat_mul(d.rot_l2g, E_local);
}

// interpolate_field (fieldArrayLoader.cpp:103-162): linear axis search with
// clamping outside the grid, x-fastest index i + nx*(j + ny*k).
void locate(const double* arr, uint32_t n, double v, size_t& i0, size_t& i1, double& alpha) {
    if (v <= arr[0]) {
        i0 = 0; i1 = 1; alpha = 0.0;
    } else if (v >= arr[n - 1]) {
        i0 = n - 2; i1 = n - 1; alpha = 1.0;
    } else {
        for (size_t i = 0; i < n - 1; ++i) {
            if (v >= arr[i] && v <= arr[i + 1]) {
                i0 = i;
                i1 = i + 1;
                const double dv = arr[i + 1] - arr[i];
                alpha = (dv > 0.0) ? (v - arr[i]) / dv : 0.0;
                return;
            }
        }
        i0 = n - 2; i1 = n - 1; alpha = 1.0;
    }
}

V3 interpolate_arrays(const oracle_field& f, const double* Ex, const double* Ey, const double* Ez,
                      V3 pos) {
    size_t ix0, ix1, iy0, iy1, iz0, iz1;
    double ax, ay, az;
    locate(f.xs, f.nx, pos.x, ix0, ix1, ax);
    locate(f.ys, f.ny, pos.y, iy0, iy1, ay);
    locate(f.zs, f.nz, pos.z, iz0, iz1, az);
    V3 r{0.0, 0.0, 0.0};
    for (size_t dx = 0; dx <= 1; ++dx) {
        const double wx = (dx == 0 ? (1.0 - ax) : ax);
        const size_t ix = (dx == 0 ? ix0 : ix1);
        for (size_t dy = 0; dy <= 1; ++dy) {
            const double wy = (dy == 0 ? (1.0 - ay) : ay);
            const size_t iy = (dy == 0 ? iy0 : iy1);
            for (size_t dz = 0; dz <= 1; ++dz) {
                const double wz = (dz == 0 ? (1.0 - az) : az);
                const size_t iz = (dz == 0 ? iz0 : iz1);
                const double w = wx * wy * wz;
                const size_t idx = ix + f.nx * (iy + f.ny * iz);
                r.x += w * Ex[idx];
                r.y += w * Ey[idx];
                r.z += w * Ez[idx];
            }
        }
    }
    return r;
}

// CompositeFieldProvider::compute_scale_factor (CompositeFieldProvider.h:117-147);
// ValueOrWaveform values are the domain's constants (eval_voltage, :150-172).
// CompositeFieldProvider::eval_voltage (CompositeFieldProvider.h:150-171): the
// constant, else a library waveform evaluated at t; an inline waveform is not
// looked up there and gives 0.
double composite_voltage(const oracle_field* f, int slot, double constant, double t) {
    if (!f || !f->waveforms[slot]) return constant;
    const icarion_waveform_desc& w = *f->waveforms[slot];
    if (w.source == ICARION_WAVEFORM_INLINE) return 0.0;
    return waveform_value(w, t);
}

// CompositeFieldProvider::compute_scale_factor (CompositeFieldProvider.h:117-147).
double term_scale(const icarion_domain_desc& d, const oracle_field_term& term, double t,
                  const oracle_field* f = nullptr) {
    switch (term.scale_kind) {
        case ICARION_FIELD_SCALE_CONSTANT: return term.constant_scale;
        case ICARION_FIELD_SCALE_DC_AXIAL: return composite_voltage(f, ICARION_SLOT_DC_AXIAL, d.dc_axial_V, t);
        case ICARION_FIELD_SCALE_DC_QUAD: return composite_voltage(f, ICARION_SLOT_DC_QUAD, d.dc_quad_V, t);
        case ICARION_FIELD_SCALE_DC_RADIAL: return composite_voltage(f, ICARION_SLOT_DC_RADIAL, d.dc_radial_V, t);
        case ICARION_FIELD_SCALE_RF: {
            const double V = composite_voltage(f, ICARION_SLOT_RF_VOLTAGE, d.rf_voltage_V, t);
            const double fr = term.frequency_Hz > 0.0
                                  ? term.frequency_Hz
                                  : composite_voltage(f, ICARION_SLOT_RF_FREQUENCY, d.rf_frequency_Hz, t);
            const double omega = 2.0 * kPi * fr;
            return V * std::cos(omega * t + term.phase_rad);
        }
        default: return 0.0;
    }
}

// The field source of a domain with field arrays: GridFieldProvider for one
// array, CompositeFieldProvider::get_E (CompositeFieldProvider.h:80-97) for
// several (per term: interpolate, then add scale * E_i in term order).
V3 interpolate(const oracle_field& f, V3 pos, const icarion_domain_desc* d = nullptr, double t = 0.0) {
    if (f.n_terms == 0) return interpolate_arrays(f, f.Ex, f.Ey, f.Ez, pos);
    if (f.n_terms == 1) return interpolate_arrays(f, f.terms[0].Ex, f.terms[0].Ey, f.terms[0].Ez, pos);
    V3 E{0.0, 0.0, 0.0};
    for (uint32_t i = 0; i < f.n_terms; ++i) {
        const oracle_field_term& term = f.terms[i];
        const double scale = d ? term_scale(*d, term, t, &f) : 1.0;
        const V3 Ei = interpolate_arrays(f, term.Ex, term.Ey, term.Ez, pos);
        E.x += scale * Ei.x;
        E.y += scale * Ei.y;
        E.z += scale * Ei.z;
    }
    return E;
}

// ElectricFieldForce::compute_soa -> IFieldModel::E (ElectricFieldForce.cpp:66-83);
// a field array replaces the analytic model (PhysicsSetup.cpp:107-195).
V3 field_E(const icarion_domain_desc& d, const oracle_field* f, V3 pos, double t) {
    if (d.use_field_array && f) {
        V3 E = interpolate(*f, pos, &d, t);
        if (d.field_array_scale != 1.0) E = E * d.field_array_scale;
        return E;
    }
    switch (d.instrument) {
        case ICARION_INSTRUMENT_IMS:
        case ICARION_INSTRUMENT_TIMS:
            return ims_field(d, pos, t, f);
        case ICARION_INSTRUMENT_QUADRUPOLE:
            return quadrupole_field(d, pos, t, f);
        case ICARION_INSTRUMENT_ORBITRAP:
            return orbitrap_field(d, pos, f);
        default:
            return V3{0.0, 0.0, 0.0};
    }
}

// ---- geometry -------------------------------------------------------------

// OrbitrapGeometry::orbitrap_surface_residual (OrbitrapGeometry.h:173-178)
double orbitrap_residual(double r, double z, double R, double R_m) {
    const double term1 = z * z;
    const double term2 = 0.5 * (r * r - R * R);
    const double term3 = R_m * R_m * std::log(R / r);
    return term1 - term2 - term3;
}

// Ions per OpenMP work chunk: 64, or fewer for small ensembles so that every
// thread gets work (long runs of a few hundred ions).
long long omp_chunk(size_t n) {
#ifdef _OPENMP
    const long long per = static_cast<long long>(n) / (4LL * omp_get_max_threads());
#else
    const long long per = static_cast<long long>(n);
#endif
    return std::max(1LL, std::min(64LL, per));
}

// OrbitrapGeometry::orbitrap_r_for_z (OrbitrapGeometry.h:181-219)
double orbitrap_r_for_z(double z, double R, double R_m) {
    const double z_abs = std::fabs(z);
    const double eps = 1e-10;
    const int max_iter = 80;
    if (z_abs < eps) return R;
    double r_lo = 0.1 * R;
    double r_hi = R;
    double f_lo = orbitrap_residual(r_lo, z_abs, R, R_m);
    double f_hi = orbitrap_residual(r_hi, z_abs, R, R_m);
    int expand_iter = 0;
    while (f_lo * f_hi > 0.0 && expand_iter < 10) {
        r_lo *= 0.5;
        r_hi *= 1.5;
        f_lo = orbitrap_residual(r_lo, z_abs, R, R_m);
        f_hi = orbitrap_residual(r_hi, z_abs, R, R_m);
        expand_iter++;
    }
    double r_mid = R;
    for (int i = 0; i < max_iter; ++i) {
        r_mid = 0.5 * (r_lo + r_hi);
        const double f_mid = orbitrap_residual(r_mid, z_abs, R, R_m);
        if (f_lo * f_mid <= 0.0) {
            r_hi = r_mid;
            f_hi = f_mid;
        } else {
            r_lo = r_mid;
            f_lo = f_mid;
        }
        if (std::fabs(f_mid) < eps) break;
    }
    return r_mid;
}

// CylindricalGeometry::contains (CylindricalGeometry.h:35-40) and
// OrbitrapGeometry::contains (OrbitrapGeometry.h:36-51); single domain, so
// DomainManager::find_domain_index (DomainManager.cpp:66-109) reduces to it.
bool contains(const icarion_domain_desc& d, V3 gpos) {
    const V3 local = mat_mul(d.rot_g2l, gpos - V3{d.origin_m[0], d.origin_m[1], d.origin_m[2]});
    if (d.instrument == ICARION_INSTRUMENT_ORBITRAP) {
        const double z_abs = std::fabs(local.z);
        const double z_max = 0.5 * d.length_m;
        if (z_abs > z_max + kGeomEpsilon) return false;
        const double r2 = local.x * local.x + local.y * local.y;
        const double r = std::sqrt(r2);
        const double r_in_allowed = orbitrap_r_for_z(z_abs, d.radius_in_m, d.radius_char_m);
        const double r_out_allowed = orbitrap_r_for_z(z_abs, d.radius_out_m, d.radius_char_m);
        if (!(r_in_allowed > 0.0 && r_out_allowed > r_in_allowed)) return false;
        return (r >= r_in_allowed - kGeomEpsilon) && (r <= r_out_allowed + kGeomEpsilon);
    }
    if (local.z < -kGeomEpsilon || local.z >= d.length_m + kGeomEpsilon) return false;
    const double r2 = local.x * local.x + local.y * local.y;
    return r2 <= (d.radius_m + kGeomEpsilon) * (d.radius_m + kGeomEpsilon);
}

// ---- damping (DampingForce.cpp:105-277) -----------------------------------
double damping_gamma(const icarion_domain_desc& d, int model, const icarion_species_desc& sp,
                     double m_ion, double q, V3 vel) {
    const double gas_density = d.gas_density_m3;
    const double temperature_K = d.temperature_K;
    const double m_neutral = d.gas_mass_kg;
    const double v_th = (temperature_K > 0.0 && m_neutral > 0.0)
        ? std::sqrt(8.0 * kBoltzmann * temperature_K / (kPi * m_neutral))
        : 0.0;
    if (model == ICARION_COLLISION_LANGEVIN) {
        // DampingForce.cpp:150-177
        const double alpha = d.gas_polarizability_m3;
        if (alpha <= 0.0 || gas_density <= 0.0 || v_th <= 0.0 || m_neutral <= 0.0) return 0.0;
        const double v_mag = std::max(1e-6, norm(vel));
        const double m_reduced = (m_ion * m_neutral) / (m_ion + m_neutral);
        const double epsilon_factor = 4.0 * kPi * kEpsilon0;
        const double cs = kPi * q * std::sqrt(alpha / (epsilon_factor * m_reduced)) / v_mag;
        return gas_density * cs * v_th * m_neutral / (m_neutral + m_ion);
    }
    if (model == ICARION_COLLISION_FRICTION) {
        // DampingForce.cpp:179-272, single-gas branch
        const double K0_cm2_Vs = sp.K0_cm2_Vs;
        if (K0_cm2_Vs <= 0.0 || m_ion <= 0.0 || gas_density <= 0.0) return 0.0;
        const double K0_m2_Vs = K0_cm2_Vs * 1e-4;
        const double N_ratio = kLoschmidt / gas_density;
        const double teff_scale = std::sqrt(kStpTemp / std::max(temperature_K, 1.0));
        const double K0_teff_m2_Vs = K0_m2_Vs * teff_scale;
        const double ion_mobility = K0_teff_m2_Vs * N_ratio;
        return q / (ion_mobility * m_ion);
    }
    return 0.0;
}

// EnvironmentConfig::gas_velocity_at (EnvironmentConfig.h:96-129) with the
// domain's geometry and rotations, as the engine passes them
// (SimulationEngineCollisionRuntime.cpp:93-101, 122-130; DampingForce.cpp:
// 70-88): Constant (and AxialUniform, given as a constant) the descriptor's
// gas velocity; AxialParabolic R_l2g (0, 0, u_max max(0, 1 - r^2/R^2)) with r
// from R_g2l (pos - origin), u_max everywhere when radius_m <= 0.
V3 gas_velocity_at(const icarion_domain_desc& d, const oracle_field* f, V3 pos) {
    if (!f || f->flow_model != ICARION_FLOW_AXIAL_PARABOLIC)
        return V3{d.gas_velocity_m_s[0], d.gas_velocity_m_s[1], d.gas_velocity_m_s[2]};
    const double u_max = f->flow_max_velocity_m_s;
    double uz = u_max;
    if (d.radius_m > 0.0) {
        const V3 rel{pos.x - d.origin_m[0], pos.y - d.origin_m[1], pos.z - d.origin_m[2]};
        const double* G = d.rot_g2l;
        const V3 local{G[0] * rel.x + G[1] * rel.y + G[2] * rel.z, G[3] * rel.x + G[4] * rel.y + G[5] * rel.z,
                       G[6] * rel.x + G[7] * rel.y + G[8] * rel.z};
        const double r2 = local.x * local.x + local.y * local.y;
        const double R2 = d.radius_m * d.radius_m;
        const double rr = R2 > 0.0 ? r2 / R2 : 0.0;
        uz = u_max * std::max(0.0, 1.0 - rr);
    }
    const double* L = d.rot_l2g;
    return V3{L[0] * 0.0 + L[1] * 0.0 + L[2] * uz, L[3] * 0.0 + L[4] * 0.0 + L[5] * uz,
              L[6] * 0.0 + L[7] * 0.0 + L[8] * uz};
}

// ForceRegistry::compute_total_force_soa (ForceRegistry.cpp:42-88):
// ElectricFieldForce (F = E q) + DampingForce (F = (v - u(pos)) * (-gamma m)).
V3 total_force(const icarion_domain_desc& d, const oracle_field* f, int coll,
               const icarion_species_desc& sp, double m, double q, V3 pos, V3 vel, double t) {
    V3 F = field_E(d, f, pos, t) * q;
    if (coll == ICARION_COLLISION_LANGEVIN || coll == ICARION_COLLISION_FRICTION) {
        const double gamma = damping_gamma(d, coll, sp, m, q, vel);
        if (gamma > 0.0) {
            const V3 u = gas_velocity_at(d, f, pos);
            F = F + (vel - u) * (-gamma * m);
        }
    }
    return F;
}

// ---- pushes ------------------------------------------------------------

// RK4Strategy::step_batch_with_stage_refresh (RK4Strategy.cpp:181-361):
// stage times t, t+h/2, t+h/2, t+h; a = F * (1/m).
void rk4_step(const icarion_domain_desc& d, const oracle_field* f, int coll,
              const icarion_species_desc& sp, double m, double q, V3& pos, V3& vel,
              double t, double dt) {
    const double half_dt = dt * 0.5;
    const double inv_m = 1.0 / m;
    const V3 p0 = pos, v0 = vel;
    const V3 k1v = v0;
    const V3 k1a = total_force(d, f, coll, sp, m, q, p0, v0, t) * inv_m;
    const V3 p2 = p0 + k1v * half_dt, v2 = v0 + k1a * half_dt;
    const V3 k2v = v2;
    const V3 k2a = total_force(d, f, coll, sp, m, q, p2, v2, t + half_dt) * inv_m;
    const V3 p3 = p0 + k2v * half_dt, v3 = v0 + k2a * half_dt;
    const V3 k3v = v3;
    const V3 k3a = total_force(d, f, coll, sp, m, q, p3, v3, t + half_dt) * inv_m;
    const V3 p4 = p0 + k3v * dt, v4 = v0 + k3a * dt;
    const V3 k4v = v4;
    const V3 k4a = total_force(d, f, coll, sp, m, q, p4, v4, t + dt) * inv_m;
    pos = p0 + (k1v + k2v * 2.0 + k3v * 2.0 + k4v) * (dt / 6.0);
    vel = v0 + (k1a + k2a * 2.0 + k3a * 2.0 + k4a) * (dt / 6.0);
}

// BorisStrategy::step (BorisStrategy.cpp:18-98): the half kicks use the total
// registry force at the step start (damping included); x += v+ dt with the
// pre-second-kick velocity.
void boris_step(const icarion_domain_desc& d, const oracle_field* f, int coll,
                const icarion_species_desc& sp, double m, double q, V3& pos, V3& vel,
                double t, double dt) {
    const double qm = q / m;
    const V3 F = total_force(d, f, coll, sp, m, q, pos, vel, t);
    const V3 a = F / m;
    V3 B{0, 0, 0};
    if (d.magnetic_enabled) B = V3{d.B_T[0], d.B_T[1], d.B_T[2]};
    V3 v_minus = vel;
    v_minus = v_minus + a * (dt * 0.5);
    const V3 t_vec = B * (qm * dt * 0.5);
    const double t_mag_sq = t_vec.x * t_vec.x + t_vec.y * t_vec.y + t_vec.z * t_vec.z;
    const double s_factor = 2.0 / (1.0 + t_mag_sq);
    const V3 s_vec = t_vec * s_factor;
    const V3 vmxt{v_minus.y * t_vec.z - v_minus.z * t_vec.y,
                  v_minus.z * t_vec.x - v_minus.x * t_vec.z,
                  v_minus.x * t_vec.y - v_minus.y * t_vec.x};
    const V3 v_prime = v_minus + vmxt;
    const V3 vpxs{v_prime.y * s_vec.z - v_prime.z * s_vec.y,
                  v_prime.z * s_vec.x - v_prime.x * s_vec.z,
                  v_prime.x * s_vec.y - v_prime.y * s_vec.x};
    const V3 v_plus = v_minus + vpxs;
    pos = pos + v_plus * dt;
    vel = v_plus + a * (dt * 0.5);
}

// ---- RK45 (RK45Strategy.cpp:101-520) ---------------------------------------
// Dormand-Prince 5(4) tableau (Hairer, Norsett, Wanner 1993; the reference's
// RK45Strategy.cpp:101-157).  The reference calls the propagated solution (the
// a7 row, FSAL) "y4" and the embedded solution "y5"; the names are kept.
namespace dp {
constexpr double c2 = 1.0 / 5.0, c3 = 3.0 / 10.0, c4 = 4.0 / 5.0, c5 = 8.0 / 9.0, c6 = 1.0,
                 c7 = 1.0;
constexpr double a21 = 1.0 / 5.0;
constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0,
                 a54 = -212.0 / 729.0;
constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                 a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0;
constexpr double a71 = 35.0 / 384.0, a72 = 0.0, a73 = 500.0 / 1113.0, a74 = 125.0 / 192.0,
                 a75 = -2187.0 / 6784.0, a76 = 11.0 / 84.0;
constexpr double b51 = 5179.0 / 57600.0, b52 = 0.0, b53 = 7571.0 / 16695.0, b54 = 393.0 / 640.0,
                 b55 = -92097.0 / 339200.0, b56 = 187.0 / 2100.0, b57 = 1.0 / 40.0;
// PI controller (RK45Strategy.cpp:159-171)
constexpr double kBeta = 0.04;
constexpr double kAlpha = 0.2 - kBeta * 0.75;
constexpr double kRejectExp = 0.2;
constexpr int kMaxReject = 10;
constexpr double kMinError = 1e-10;
constexpr double kDtMinTol = 1.001;
}  // namespace dp

// RK45Strategy::estimate_error (RK45Strategy.cpp:272-313)
double rk45_error(const icarion_rk45_desc& c, const double y4[6], const double y5[6],
                  const double y0[6]) {
    double m = 0.0;
    for (int i = 0; i < 6; ++i) {
        const double e = std::fabs(y5[i] - y4[i]);
        const double sc = c.atol + c.rtol * std::fabs(y0[i]);
        m = std::max(m, e / sc);
    }
    return std::max(m, dp::kMinError);
}

// RK45Strategy::compute_new_step (RK45Strategy.cpp:315-351)
double rk45_new_step(const icarion_rk45_desc& c, double dt, double err, double dt_min,
                     double dt_max, double prev_err) {
    double f;
    if (err > 1.0) {
        f = c.safety_factor * std::pow(1.0 / err, dp::kRejectExp);
    } else {
        f = c.safety_factor * std::pow(prev_err / err, dp::kBeta) * std::pow(1.0 / err, dp::kAlpha);
    }
    f = std::max(f, c.max_step_decrease);
    f = std::min(f, c.max_step_increase);
    double dn = dt * f;
    dn = std::max(dn, dt_min);
    dn = std::min(dn, dt_max);
    return dn;
}

// RK45Strategy::step (RK45Strategy.cpp:354-520) for one ion from time t with
// trial dt.  Returns the accepted dt (-1 when no trial was accepted).
double rk45_step(const icarion_domain_desc& d, const oracle_field* f, int coll,
                 const icarion_rk45_desc& cfg, const icarion_species_desc& sp, double m, double q,
                 V3& pos, V3& vel, double t, double dt, oracle_rk45_ion& st, double& dt_next,
                 int& rejected) {
    const double inv_m = 1.0 / m;
    auto acc = [&](V3 p, V3 v, double tt) { return total_force(d, f, coll, sp, m, q, p, v, tt) * inv_m; };
    const double dt_min = std::max(dt * cfg.min_step_factor, cfg.absolute_min_step_s);
    const double dt_max = dt * cfg.max_step_factor;
    double h = dt;
    rejected = 0;
    const V3 p0 = pos, v0 = vel;
    for (int attempt = 0; attempt < dp::kMaxReject; ++attempt) {
        using namespace dp;
        const V3 k1v = v0;
        const V3 k1a = st.fsal ? V3{st.k1a[0], st.k1a[1], st.k1a[2]} : acc(p0, v0, t);
        const V3 p2 = p0 + k1v * (h * a21), v2 = v0 + k1a * (h * a21);
        const V3 k2v = v2, k2a = acc(p2, v2, t + c2 * h);
        const V3 p3 = p0 + (k1v * a31 + k2v * a32) * h, v3 = v0 + (k1a * a31 + k2a * a32) * h;
        const V3 k3v = v3, k3a = acc(p3, v3, t + c3 * h);
        const V3 p4 = p0 + (k1v * a41 + k2v * a42 + k3v * a43) * h;
        const V3 v4 = v0 + (k1a * a41 + k2a * a42 + k3a * a43) * h;
        const V3 k4v = v4, k4a = acc(p4, v4, t + c4 * h);
        const V3 p5 = p0 + (k1v * a51 + k2v * a52 + k3v * a53 + k4v * a54) * h;
        const V3 v5 = v0 + (k1a * a51 + k2a * a52 + k3a * a53 + k4a * a54) * h;
        const V3 k5v = v5, k5a = acc(p5, v5, t + c5 * h);
        const V3 p6 = p0 + (k1v * a61 + k2v * a62 + k3v * a63 + k4v * a64 + k5v * a65) * h;
        const V3 v6 = v0 + (k1a * a61 + k2a * a62 + k3a * a63 + k4a * a64 + k5a * a65) * h;
        const V3 k6v = v6, k6a = acc(p6, v6, t + c6 * h);
        const V3 p7 = p0 + (k1v * a71 + k2v * a72 + k3v * a73 + k4v * a74 + k5v * a75 + k6v * a76) * h;
        const V3 v7 = v0 + (k1a * a71 + k2a * a72 + k3a * a73 + k4a * a74 + k5a * a75 + k6a * a76) * h;
        const V3 k7v = v7, k7a = acc(p7, v7, t + c7 * h);
        const V3 p5s = p0 + (k1v * b51 + k2v * b52 + k3v * b53 + k4v * b54 + k5v * b55 +
                             k6v * b56 + k7v * b57) * h;
        const V3 v5s = v0 + (k1a * b51 + k2a * b52 + k3a * b53 + k4a * b54 + k5a * b55 +
                             k6a * b56 + k7a * b57) * h;
        const double y4[6] = {p7.x, p7.y, p7.z, v7.x, v7.y, v7.z};
        const double y5[6] = {p5s.x, p5s.y, p5s.z, v5s.x, v5s.y, v5s.z};
        const double y0[6] = {p0.x, p0.y, p0.z, v0.x, v0.y, v0.z};
        const double err = rk45_error(cfg, y4, y5, y0);
        const bool force_accept = cfg.accept_at_dt_min && (h <= dt_min * kDtMinTol);
        if (err <= 1.0 || force_accept) {
            pos = p7;
            vel = v7;
            st.k1a[0] = k7a.x; st.k1a[1] = k7a.y; st.k1a[2] = k7a.z;
            st.fsal = 1;
            dt_next = rk45_new_step(cfg, h, err, dt_min, dt_max, st.last_error);
            st.last_error = err;
            return h;
        }
        h = rk45_new_step(cfg, h, err, dt_min, dt_max, st.last_error);
        ++rejected;
        st.fsal = 0;
    }
    return -1.0;
}

// ---- collision kernels (CollisionKernels.cpp, VelocitySampling.cpp,
// CollisionGeometry.cpp) ------------------------------------------------------

// VelocitySampling::box_muller_sample (VelocitySampling.cpp:56-66)
template <class Rng>
double box_muller(double sigma, Rng& rng) {
    const double u1 = rng.uniform01();
    const double u2 = rng.uniform01();
    return sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

// VelocitySampling::sample_neutral_velocity (VelocitySampling.cpp:9-35)
template <class Rng>
V3 sample_neutral_velocity(double T_K, double mass_kg, V3 flow, Rng& rng) {
    const double sigma = std::sqrt(kBoltzmann * T_K / mass_kg);
    const double vx = box_muller(sigma, rng);
    const double vy = box_muller(sigma, rng);
    const double vz = box_muller(sigma, rng);
    return V3{vx + flow.x, vy + flow.y, vz + flow.z};
}

// CollisionKernels::sample_isotropic_direction (CollisionKernels.cpp:383-397)
template <class Rng>
V3 sample_isotropic_direction(Rng& rng) {
    const double u1 = rng.uniform01();
    const double u2 = rng.uniform01();
    const double cosT = 2.0 * u1 - 1.0;
    const double sinT = std::sqrt(std::max(0.0, 1.0 - cosT * cosT));
    const double phi = 2.0 * kPi * u2;
    return V3{sinT * std::cos(phi), sinT * std::sin(phi), cosT};
}

// CollisionKernels::hss_collision (CollisionKernels.cpp:139-169)
template <class Rng>
V3 hss_collision(V3 v_ion, V3 v_neutral, double ion_mass, double neutral_mass, Rng& rng) {
    V3 vrel = v_ion - v_neutral;
    double vrel_mag = norm(vrel);
    if (vrel_mag <= kMinVelocityMag) {
        vrel = sample_isotropic_direction(rng) * kMinVelocityMag;
        vrel_mag = kMinVelocityMag;
    }
    const V3 new_dir = sample_isotropic_direction(rng);
    const V3 vrel_scattered = new_dir * vrel_mag;
    const double mt = ion_mass + neutral_mass;
    const double inv_mt = 1.0 / mt;
    const V3 Vcom = (v_ion * ion_mass + v_neutral * neutral_mass) * inv_mt;
    return Vcom + vrel_scattered * (neutral_mass * inv_mt);
}

// CollisionGeometry::construct_orthonormal_basis (CollisionGeometry.cpp:11-28)
void orthonormal_basis(V3 ehat, V3& t1, V3& t2) {
    const V3 a = (std::fabs(ehat.x) < 0.9) ? V3{1, 0, 0} : V3{0, 1, 0};
    t1 = normalize(cross(ehat, a));
    t2 = normalize(cross(ehat, t1));
}

// CollisionGeometry::generate_random_rotation (CollisionGeometry.cpp:30-59),
// Shoemake uniform quaternion; R row-major 3x3.
template <class Rng>
void random_rotation(Rng& rng, double R[3][3]) {
    const double u1 = rng.uniform01();
    const double u2 = rng.uniform01();
    const double u3 = rng.uniform01();
    const double q1 = std::sqrt(1.0 - u1) * std::sin(2.0 * kPi * u2);
    const double q2 = std::sqrt(1.0 - u1) * std::cos(2.0 * kPi * u2);
    const double q3 = std::sqrt(u1) * std::sin(2.0 * kPi * u3);
    const double q4 = std::sqrt(u1) * std::cos(2.0 * kPi * u3);
    const double x = q1, y = q2, z = q3, w = q4;
    R[0][0] = 1 - 2 * (y * y + z * z);
    R[0][1] = 2 * (x * y - z * w);
    R[0][2] = 2 * (x * z + y * w);
    R[1][0] = 2 * (x * y + z * w);
    R[1][1] = 1 - 2 * (x * x + z * z);
    R[1][2] = 2 * (y * z - x * w);
    R[2][0] = 2 * (x * z - y * w);
    R[2][1] = 2 * (y * z + x * w);
    R[2][2] = 1 - 2 * (x * x + y * y);
}

// CollisionKernels::ehss_collision_with_orientation + rotate_and_analyze_molecule
// + sample_impact_geometry + detect_atom_collision + compute_reflected_velocity
// + to_lab_frame (CollisionKernels.cpp:44-137, 241-381): the molecule in
// orientation orient (collision-axis frame), the impact range widened to
// sqrt(sigma_eff / pi) when larger, up to max_attempts impact parameters.
template <class Rng>
V3 ehss_collision_oriented(V3 v_ion, V3 v_neutral, double ion_mass, double neutral_mass,
                           double ion_radius, const double* atoms, uint32_t nat, const double orient[3][3],
                           double sigma_eff, int max_attempts, Rng& rng, std::vector<V3>& rotated) {
    if (nat == 0) return v_ion;
    V3 vrel = v_ion - v_neutral;
    double vrel_mag = norm(vrel);
    if (vrel_mag <= kMinVelocityMag) {
        vrel = sample_isotropic_direction(rng) * kMinVelocityMag;
        vrel_mag = kMinVelocityMag;
    }
    const V3 axis = vrel / vrel_mag;
    V3 t1, t2;
    orthonormal_basis(axis, t1, t2);
    const double basis[3][3] = {{t1.x, t2.x, axis.x}, {t1.y, t2.y, axis.y}, {t1.z, t2.z, axis.z}};
    double Rm[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Rm[i][j] = basis[i][0] * orient[0][j] + basis[i][1] * orient[1][j] +
                       basis[i][2] * orient[2][j];
    rotated.resize(nat);
    double b_max = 0.0;
    for (uint32_t j = 0; j < nat; ++j) {
        const V3 c{atoms[4 * j], atoms[4 * j + 1], atoms[4 * j + 2]};
        const V3 ra{Rm[0][0] * c.x + Rm[0][1] * c.y + Rm[0][2] * c.z,
                    Rm[1][0] * c.x + Rm[1][1] * c.y + Rm[1][2] * c.z,
                    Rm[2][0] * c.x + Rm[2][1] * c.y + Rm[2][2] * c.z};
        rotated[j] = ra;
        const double sra = dot(axis, ra);
        const V3 ra_perp = ra - axis * sra;
        const double cur = norm(ra_perp) + atoms[4 * j + 3];
        if (cur > b_max) b_max = cur;
    }
    b_max += ion_radius;
    if (sigma_eff > 0.0) {
        const double b_sigma = std::sqrt(sigma_eff / kPi);
        if (b_sigma > b_max) b_max = b_sigma;
    }
    for (int attempt = 0; attempt < max_attempts; ++attempt) {
        V3 s1, s2;
        orthonormal_basis(axis, s1, s2);
        const double uA = rng.uniform01();
        const double uB = rng.uniform01();
        const double b = b_max * std::sqrt(uA);
        const double phi = 2.0 * kPi * uB;
        const V3 offset = s1 * (b * std::cos(phi)) + s2 * (b * std::sin(phi));
        bool hit = false;
        V3 n_best{0, 0, 0};
        double best_s = 0.0;
        for (uint32_t j = 0; j < nat; ++j) {
            const V3 ra = rotated[j];
            const double Rsum = atoms[4 * j + 3] + ion_radius;
            const double Rsum2 = Rsum * Rsum;
            const V3 rel = offset - ra;
            const double sstar = dot(axis, rel);
            const V3 dvec = rel - axis * sstar;
            const double dmin2 = dot(dvec, dvec);
            if (dmin2 <= Rsum2) {
                const double h = std::sqrt(std::max(0.0, Rsum2 - dmin2));
                const double s_hit = sstar - h;
                const V3 p_hit_rel = rel - axis * s_hit;
                const double phr2 = dot(p_hit_rel, p_hit_rel);
                V3 n_contact;
                if (phr2 <= kMinContactDistSq) {
                    n_contact = axis * -1.0;
                } else {
                    n_contact = p_hit_rel * (1.0 / std::sqrt(phr2));
                    if (dot(axis, n_contact) > 0.0) n_contact = n_contact * -1.0;
                }
                if (!hit || s_hit < best_s) {
                    hit = true;
                    n_best = n_contact;
                    best_s = s_hit;
                }
            }
        }
        if (hit) {
            const double vdotn = dot(vrel, n_best);
            const V3 vrel_reflected = vrel - n_best * (2.0 * vdotn);
            const double mt = ion_mass + neutral_mass;
            const double inv_mt = 1.0 / mt;
            const V3 Vcom = (v_ion * ion_mass + v_neutral * neutral_mass) * inv_mt;
            return Vcom + vrel_reflected * (neutral_mass * inv_mt);
        }
        if (attempt == kBmaxExpansionAttempt) b_max *= kBmaxExpansionFactor;
    }
    return v_ion;
}

// CollisionKernels::ehss_collision (CollisionKernels.cpp:13-42): a random
// orientation first, then the oriented kernel with sigma_eff_m2 = 0 and
// max_attempts = 256 as in the handler's online call (EHSSCollisionHandler.cpp:294-296).
template <class Rng>
V3 ehss_collision(V3 v_ion, V3 v_neutral, double ion_mass, double neutral_mass,
                  double ion_radius, const double* atoms, uint32_t nat, Rng& rng,
                  std::vector<V3>& rotated) {
    double orient[3][3];
    random_rotation(rng, orient);
    return ehss_collision_oriented(v_ion, v_neutral, ion_mass, neutral_mass, ion_radius, atoms, nat, orient,
                                   0.0, kEhssMaxAttempts, rng, rotated);
}

// CollisionGeometry::quaternion_to_rotation (CollisionGeometry.cpp:73-104).
void quaternion_to_rotation(const double q[4], double R[3][3]) {
    double w = q[0], x = q[1], y = q[2], z = q[3];
    const double nrm = std::sqrt(w * w + x * x + y * y + z * z);
    if (nrm <= 0.0) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) R[i][j] = (i == j) ? 1.0 : 0.0;
        return;
    }
    w /= nrm;
    x /= nrm;
    y /= nrm;
    z /= nrm;
    R[0][0] = 1.0 - 2.0 * (y * y + z * z);
    R[0][1] = 2.0 * (x * y - z * w);
    R[0][2] = 2.0 * (x * z + y * w);
    R[1][0] = 2.0 * (x * y + z * w);
    R[1][1] = 1.0 - 2.0 * (x * x + z * z);
    R[1][2] = 2.0 * (y * z - x * w);
    R[2][0] = 2.0 * (x * z - y * w);
    R[2][1] = 2.0 * (y * z + x * w);
    R[2][2] = 1.0 - 2.0 * (x * x + y * y);
}

// The collision tables one ion sees: its row of the mixture cross sections
// and its species' offline EHSS sample set.
struct IonTables {
    const oracle_gas_mixture* mix = nullptr;
    const double* mix_row = nullptr;
    uint32_t off_n = 0, off_stride = 0;
    const double* off_sigma = nullptr;
    const double* off_mu = nullptr;
    uint32_t ori_n = 0;                 // orientation samples of the species
    const double* ori_quat = nullptr;   // 4 per orientation (w, x, y, z)
    const double* ori_area = nullptr;   // the area for the domain's gas
};

// EHSSCollisionHandler::handle_collision with orientation samples, single gas
// (EHSSCollisionHandler.cpp:187-217, 219-313, 426-455): the orientation drawn
// first, its area as sigma_eff in the test; on a collision the oriented
// first-contact kernel with that sigma_eff, 256 attempts and force_hit (x16).
template <class Rng>
bool ehss_orientation_collide(const icarion_domain_desc& d, uint32_t n_orient, const double* quat,
                              const double* area, const double* atoms, uint32_t nat, double m_ion, V3& vel,
                              double dt, Rng& rng, std::vector<V3>& scratch, uint32_t* orient = nullptr) {
    size_t idx = static_cast<size_t>(rng.uniform01() * static_cast<double>(n_orient));
    if (idx >= n_orient) idx = n_orient - 1;
    if (orient) *orient = static_cast<uint32_t>(idx);
    const double sigma_eff = area[idx];
    if (sigma_eff <= 0.0) return false;
    const double m_neutral = d.gas_mass_kg;
    const V3 v_gas{d.gas_velocity_m_s[0], d.gas_velocity_m_s[1], d.gas_velocity_m_s[2]};
    const V3 v_neutral = sample_neutral_velocity(d.temperature_K, m_neutral, v_gas, rng);
    const V3 v_rel = vel - v_neutral;
    const double v_rel_mag = norm(v_rel);
    if (v_rel_mag < kMinRelativeVelocity) return false;
    const double P = 1.0 - std::exp(-d.gas_density_m3 * sigma_eff * v_rel_mag * dt);
    if (rng.uniform01() >= P) return false;
    double R[3][3];
    quaternion_to_rotation(quat + 4 * idx, R);
    vel = ehss_collision_oriented(vel, v_neutral, m_ion, m_neutral, d.gas_radius_m, atoms, nat, R, sigma_eff,
                                  kEhssMaxAttempts * 16, rng, scratch);
    return true;
}

// EHSSCollisionHandler::handle_collision with an offline sample set, single gas
// (EHSSCollisionHandler.cpp:151-172, 222-299, 426-455).  Returns true on a
// collision; orient / mu_used report the draws.
template <class Rng>
bool ehss_offline_collide(const icarion_domain_desc& d, uint32_t n_orient, uint32_t stride,
                          const double* sig, const double* mus, double m_ion, V3& vel, double dt,
                          Rng& rng, uint32_t* orient = nullptr, double* mu_used = nullptr,
                          V3* vn_out = nullptr) {
    size_t idx = static_cast<size_t>(rng.uniform01() * static_cast<double>(n_orient));
    if (idx >= n_orient) idx = n_orient - 1;
    if (orient) *orient = static_cast<uint32_t>(idx);
    const double sigma_eff = sig[idx];
    if (sigma_eff <= 0.0) return false;
    const double m_neutral = d.gas_mass_kg;
    const V3 v_gas{d.gas_velocity_m_s[0], d.gas_velocity_m_s[1], d.gas_velocity_m_s[2]};
    const V3 v_neutral = sample_neutral_velocity(d.temperature_K, m_neutral, v_gas, rng);
    if (vn_out) *vn_out = v_neutral;
    const V3 v_rel = vel - v_neutral;
    const double v_rel_mag = norm(v_rel);
    if (v_rel_mag < kMinRelativeVelocity) return false;
    const double P = 1.0 - std::exp(-d.gas_density_m3 * sigma_eff * v_rel_mag * dt);
    if (rng.uniform01() >= P) return false;
    size_t mu_idx = static_cast<size_t>(rng.uniform01() * static_cast<double>(stride));
    if (mu_idx >= stride) mu_idx = stride - 1;
    double mu = mus[idx * stride + mu_idx];
    mu = std::min(1.0, std::max(-1.0, mu));
    if (mu_used) *mu_used = mu;
    const double sin_theta = std::sqrt(std::max(0.0, 1.0 - mu * mu));
    const double phi = 2.0 * kPi * rng.uniform01();
    const V3 axis{v_rel.x / v_rel_mag, v_rel.y / v_rel_mag, v_rel.z / v_rel_mag};
    V3 t1, t2;
    orthonormal_basis(axis, t1, t2);
    const V3 n_contact = t1 * (sin_theta * std::cos(phi)) + t2 * (sin_theta * std::sin(phi)) +
                         axis * (-mu);
    const double vdotn = v_rel.x * n_contact.x + v_rel.y * n_contact.y + v_rel.z * n_contact.z;
    const V3 v_rel_reflected = v_rel - n_contact * (2.0 * vdotn);
    const double mt = m_ion + m_neutral;
    const double inv_mt = 1.0 / mt;
    const V3 Vcom = (vel * m_ion + v_neutral * m_neutral) * inv_mt;
    vel = Vcom + v_rel_reflected * (m_neutral * inv_mt);
    return true;
}

// HSSCollisionHandler::handle_collision mixture path (HSSCollisionHandler.cpp:105-249):
// a neutral per participating component with n_i > 0, rates from the actual
// relative speeds, one test on the total rate, the partner drawn with weight
// k_i, then hss_collision with the neutral sampled for it.  Returns the
// selected component index, or -1.
template <class Rng>
int collide_mixture(const icarion_domain_desc& d, const oracle_gas_mixture& mix,
                    const double* sigma_row, double m_ion, V3& vel, double dt, Rng& rng,
                    V3* vn_out = nullptr) {
    struct Data {
        uint32_t comp;
        double k;
        V3 vn;
    };
    Data data[ICARION_MAX_GAS_COMPONENTS];
    int nd = 0;
    double k_total = 0.0;
    const V3 flow{d.gas_velocity_m_s[0], d.gas_velocity_m_s[1], d.gas_velocity_m_s[2]};
    for (uint32_t i = 0; i < mix.n_components && i < ICARION_MAX_GAS_COMPONENTS; ++i) {
        const icarion_gas_component& c = mix.components[i];
        if (!c.participates) continue;
        const double n_i = c.density_m3;
        if (n_i <= 0.0) continue;
        const double sigma_i = sigma_row[i];
        const V3 v_neutral = sample_neutral_velocity(d.temperature_K, c.mass_kg, flow, rng);
        const V3 v_rel = vel - v_neutral;
        const double v_rel_mag = norm(v_rel);
        if (v_rel_mag < kMinRelativeVelocity) continue;
        const double k_i = n_i * sigma_i * v_rel_mag;
        data[nd++] = Data{i, k_i, v_neutral};
        k_total += k_i;
    }
    if (k_total <= 0.0 || nd == 0) return -1;
    double P_total = 1.0;
    if (k_total * dt <= 50.0) P_total = 1.0 - std::exp(-k_total * dt);
    if (rng.uniform01() >= P_total) return -1;
    const double r = rng.uniform01() * k_total;
    int sel = 0;
    double cum = 0.0;
    for (int j = 0; j < nd; ++j) {
        cum += data[j].k;
        if (r < cum) {
            sel = j;
            break;
        }
    }
    if (vn_out) *vn_out = data[sel].vn;
    vel = hss_collision(vel, data[sel].vn, m_ion, mix.components[data[sel].comp].mass_kg, rng);
    return static_cast<int>(data[sel].comp);
}

// One collision stage of one ion.  HSS: HSSCollisionHandler::handle_collision
// single-gas path (HSSCollisionHandler.cpp:257-376), or the mixture path when
// a mixture is given.  EHSS:
// EHSSCollisionHandler::handle_collision single-gas online path
// (EHSSCollisionHandler.cpp:222-374, 426-455).  Langevin / Friction with
// ou_thermal_kicks: CollisionKernels::ou_velocity_update with apply_damping =
// false (CollisionKernels.cpp:171-214) at the ion's own gamma.
// Returns true when a collision event happened.
template <class Rng>
bool collide(const icarion_domain_desc& d, const icarion_physics_desc& ph,
             const icarion_species_desc& sp, const double* atoms, double m_ion, double q,
             V3& vel, double dt, Rng& rng, std::vector<V3>& scratch,
             const IonTables& tb = IonTables{}) {
    const int model = ph.collision_model;
    const V3 v_gas{d.gas_velocity_m_s[0], d.gas_velocity_m_s[1], d.gas_velocity_m_s[2]};
    if (model == ICARION_COLLISION_HSS && tb.mix && tb.mix->n_components > 0)
        return collide_mixture(d, *tb.mix, tb.mix_row, m_ion, vel, dt, rng) >= 0;
    if (model == ICARION_COLLISION_EHSS && tb.mix && tb.mix->n_components > 0)
        return ehss_mixture_collide(d, *tb.mix, tb.mix_row, atoms + 4 * static_cast<size_t>(sp.atom_offset),
                                    sp.atom_count, m_ion, vel, dt, rng, scratch);
    if (model == ICARION_COLLISION_EHSS && tb.off_n > 0)
        return ehss_offline_collide(d, tb.off_n, tb.off_stride, tb.off_sigma, tb.off_mu, m_ion, vel,
                                    dt, rng);
    if (model == ICARION_COLLISION_EHSS && tb.ori_n > 0)
        return ehss_orientation_collide(d, tb.ori_n, tb.ori_quat, tb.ori_area,
                                        atoms + 4 * static_cast<size_t>(sp.atom_offset), sp.atom_count, m_ion,
                                        vel, dt, rng, scratch);
    if (model == ICARION_COLLISION_HSS || model == ICARION_COLLISION_EHSS) {
        const double n = d.gas_density_m3;
        const double T_K = d.temperature_K;
        const double m_neutral = d.gas_mass_kg;
        const double sigma = (model == ICARION_COLLISION_HSS) ? sp.sigma_hss_m2 : sp.sigma_ehss_m2;
        if (sigma <= 0.0) return false;
        const V3 v_neutral = sample_neutral_velocity(T_K, m_neutral, v_gas, rng);
        const V3 v_rel = vel - v_neutral;
        const double v_rel_mag = norm(v_rel);
        if (v_rel_mag < kMinRelativeVelocity) return false;
        const double P = 1.0 - std::exp(-n * sigma * v_rel_mag * dt);
        if (rng.uniform01() >= P) return false;
        if (model == ICARION_COLLISION_HSS) {
            vel = hss_collision(vel, v_neutral, m_ion, m_neutral, rng);
        } else {
            vel = ehss_collision(vel, v_neutral, m_ion, m_neutral, d.gas_radius_m,
                                 atoms + 4 * static_cast<size_t>(sp.atom_offset), sp.atom_count,
                                 rng, scratch);
        }
        return true;
    }
    if ((model == ICARION_COLLISION_LANGEVIN || model == ICARION_COLLISION_FRICTION) &&
        ph.ou_thermal_kicks) {
        const double gamma = damping_gamma(d, model, sp, m_ion, q, vel);
        if (gamma <= 0.0 || dt <= 0.0) return false;
        const double kBT_over_m = kBoltzmann * d.temperature_K / m_ion;
        const double sigma = std::sqrt(2.0 * gamma * kBT_over_m * dt);
        vel.x += sigma * rng.normal();
        vel.y += sigma * rng.normal();
        vel.z += sigma * rng.normal();
        return false;
    }
    return false;
}

// EHSSCollisionHandler::handle_collision mixture path (EHSSCollisionHandler.cpp:
// 311-424) with the online geometry method: every component in order, a
// neutral for each participating one, k_i = n_i sigma_i |v - v_i| where all
// are > 0, one test on k_total, the partner by weight (the last component
// when the running sum falls short), then apply_collision with that neutral
// and no second test (EHSSCollisionHandler.cpp:219-250, 294-296).
template <class Rng>
bool ehss_mixture_collide(const icarion_domain_desc& d, const oracle_gas_mixture& mix, const double* sig_row,
                          const double* atoms, uint32_t nat, double m_ion, V3& vel, double dt, Rng& rng,
                          std::vector<V3>& scratch) {
    const uint32_t nc = mix.n_components;
    const V3 v_gas{d.gas_velocity_m_s[0], d.gas_velocity_m_s[1], d.gas_velocity_m_s[2]};
    std::vector<double> k(nc, 0.0);
    std::vector<V3> vn(nc, V3{0.0, 0.0, 0.0});
    double k_total = 0.0;
    for (uint32_t c = 0; c < nc; ++c) {
        const icarion_gas_component& comp = mix.components[c];
        if (!comp.participates) continue;
        vn[c] = sample_neutral_velocity(d.temperature_K, comp.mass_kg, v_gas, rng);
        const double sigma_i = sig_row[c];
        const double n_i = comp.density_m3;
        const double v_rel_i = norm(vel - vn[c]);
        if (sigma_i <= 0.0 || n_i <= 0.0 || v_rel_i < kMinRelativeVelocity) continue;
        k[c] = n_i * sigma_i * v_rel_i;
        k_total += k[c];
    }
    if (k_total <= 0.0) return false;
    double P_total = 1.0;
    if (k_total * dt <= 50.0) P_total = 1.0 - std::exp(-k_total * dt);
    if (rng.uniform01() >= P_total) return false;
    const double r = rng.uniform01() * k_total;
    uint32_t idx = 0;
    double cum = 0.0;
    for (; idx < nc; ++idx) {
        cum += k[idx];
        if (r < cum) break;
    }
    if (idx >= nc) idx = nc - 1;
    const icarion_gas_component& comp = mix.components[idx];
    const double sigma_eff = sig_row[idx];
    if (sigma_eff <= 0.0) return false;
    const double radius = (mix.radius_m && mix.radius_m[idx] > 0.0) ? mix.radius_m[idx] : d.gas_radius_m;
    const V3 v_rel = vel - vn[idx];
    if (norm(v_rel) < kMinRelativeVelocity) return false;
    vel = ehss_collision(vel, vn[idx], m_ion, comp.mass_kg, radius, atoms, nat, rng, scratch);
    return true;
}

// Per-ion engine step (SimulationEngine::process_timestep,
// SimulationEngine.cpp:374-428 with prepare/finalize of
// SimulationEngineLifecycle.cpp:72-195).  The reference runs each phase over
// all ions before the next; ions do not interact, so running all phases of
// one ion back to back gives the same per-ion result.
// update_dynamic_environments (SimulationEngineLifecycle.cpp:34-47) through
// EnvironmentConfig::update_time_dependent and compute_derived_properties
// (EnvironmentConfig.h:132-175): at the engine time of each step start the
// domain's pressure takes its waveform's value, n = P / (kB T), and a gas
// mixture's components n_i = x_i n, with x_i the fractions the constant state
// holds (n_i / n of the descriptor).  ok = false: a step's P(t) <= 0, which the
// reference rejects.
struct EnvSteps {
    std::vector<icarion_domain_desc> dom;
    std::vector<oracle_gas_mixture> mix;
    std::vector<icarion_gas_component> comps;
    bool ok = true;

    EnvSteps(const icarion_domain_desc& d, const icarion_waveform_desc& pressure, const oracle_gas_mixture* m,
             const double* t_table, uint64_t n_steps) {
        dom.assign(n_steps, d);
        const uint32_t nc = m ? m->n_components : 0;
        if (nc) {
            mix.assign(n_steps, *m);
            comps.resize(n_steps * nc);
        }
        for (uint64_t k = 0; k < n_steps; ++k) {
            const double P = waveform_value(pressure, t_table[k]);
            if (!(P > 0.0)) {
                ok = false;
                return;
            }
            const double n = P / (kBoltzmann * d.temperature_K);
            dom[k].pressure_Pa = P;
            dom[k].gas_density_m3 = n;
            for (uint32_t c = 0; c < nc; ++c) {
                icarion_gas_component g = m->components[c];
                const double x = d.gas_density_m3 > 0.0 ? g.density_m3 / d.gas_density_m3 : 0.0;
                g.density_m3 = x * n;
                comps[k * nc + c] = g;
            }
            if (nc) mix[k].components = &comps[k * nc];
        }
    }
};

struct IonCtx {
    V3 pos, vel;
    double m, q;
    uint8_t active, born;
    double birth, death, arrival;
    uint32_t ncoll;
};

// rng_kind ORACLE_RNG_PHILOX: a fresh (seed, ion, step) stream per step;
// ORACLE_RNG_MT19937_64: the ion's persistent PhysicsRng (*mt).
// Collision stages per engine step: the substeps the engine's CPU path runs
// (SimulationEngineCollisionRuntime.cpp:26-33, 77-82), 1 when none are set.
int collision_substeps(const icarion_physics_desc& ph) { return std::max(1, static_cast<int>(ph.collision_substeps)); }

uint64_t run_ion(const icarion_domain_desc& d, const icarion_physics_desc& ph,
                 const icarion_species_desc& sp, const double* atoms, const oracle_field* f,
                 IonCtx& ion, const double* t_table, double dt, uint64_t n_steps,
                 uint32_t ion_index, uint64_t step0, MtRng* mt, uint64_t& collisions,
                 std::vector<V3>& scratch, const IonTables& tb, const EnvSteps* env) {
    uint64_t pushed = 0;
    const bool detect = std::isfinite(d.detector_z_m);
    for (uint64_t k = 0; k < n_steps; ++k) {
        const double t = t_table[k];
        // prepare_ions_for_integration (SimulationEngineLifecycle.cpp:86-107)
        if (!ion.born && t >= ion.birth) {
            ion.born = 1;
            ion.active = 1;
        }
        if (!ion.active || !ion.born) continue;
        if (!contains(d, ion.pos)) {
            ion.active = 0;
            ion.death = t;
            continue;
        }
        // the step's environment (a pressure waveform: its density)
        const icarion_domain_desc& de = env ? env->dom[k] : d;
        IonTables tbe = tb;
        if (env && tb.mix) tbe.mix = &env->mix[k];
        // perform_integration
        const double z_before = ion.pos.z;
        if (ph.integrator == ICARION_INTEGRATOR_BORIS)
            boris_step(de, f, ph.collision_model, sp, ion.m, ion.q, ion.pos, ion.vel, t, dt);
        else
            rk4_step(de, f, ph.collision_model, sp, ion.m, ion.q, ion.pos, ion.vel, t, dt);
        ++pushed;
        // perform_collisions: handle_collisions_cpu's substep loop
        // (SimulationEngineCollisionRuntime.cpp:84-191), collision_dt =
        // dt / substeps, each substep an event of its own, one RNG stream
        if (ph.collision_model != ICARION_COLLISION_NONE) {
            const int ns = collision_substeps(ph);
            const double collision_dt = dt / static_cast<double>(ns);
            PhiloxStream prng(ph.rng_seed, ion_index, step0 + k);
            // a non-constant flow: the environment takes the gas velocity at the
            // ion's pushed position (SimulationEngineCollisionRuntime.cpp:122-130)
            icarion_domain_desc dflow;
            const bool flow = f && f->flow_model != ICARION_FLOW_CONSTANT;
            if (flow) {
                dflow = de;
                const V3 u = gas_velocity_at(de, f, ion.pos);
                dflow.gas_velocity_m_s[0] = u.x;
                dflow.gas_velocity_m_s[1] = u.y;
                dflow.gas_velocity_m_s[2] = u.z;
            }
            const icarion_domain_desc& dc = flow ? dflow : de;
            for (int sub = 0; sub < ns; ++sub) {
                const bool hit = mt ? collide(dc, ph, sp, atoms, ion.m, ion.q, ion.vel, collision_dt, *mt, scratch, tbe)
                                    : collide(dc, ph, sp, atoms, ion.m, ion.q, ion.vel, collision_dt, prng, scratch, tbe);
                if (hit) {
                    ++ion.ncoll;
                    ++collisions;
                }
            }
        }
        // arrival: first crossing of the detector plane (linear in the step)
        if (detect && std::isnan(ion.arrival)) {
            const double z_after = ion.pos.z;
            if (z_before < d.detector_z_m && z_after >= d.detector_z_m) {
                ion.arrival = t + dt * (d.detector_z_m - z_before) / (z_after - z_before);
            }
        }
        // finalize_ions_after_integration (SimulationEngineLifecycle.cpp:137-192)
        if (!contains(d, ion.pos)) {
            ion.active = 0;
            ion.death = t;
            continue;
        }
        if (!finite3(ion.pos) || !finite3(ion.vel)) {
            ion.active = 0;
            ion.death = t;
        }
    }
    return pushed;
}

// DomainManager of a multi-domain instrument (DomainManager.cpp:16-109,
// 163-198): spans of the cylindrical domains sorted by z_min, the lookup, the
// forward axial bridge and the shared-boundary handoff.
struct Domains {
    struct Span {
        double z_min, z_max, radius;
        int idx;
    };
    const icarion_domain_desc* d;
    uint32_t n;
    const oracle_field* const* fields = nullptr;   // per domain (waveforms), may be NULL
    const oracle_gas_mixture* const* mixtures = nullptr;   // per domain (HSS gas mixture), may be NULL
    const icarion_species_desc* const* species = nullptr;  // per domain (species table), may be NULL
    std::vector<std::unique_ptr<EnvSteps>> env;            // per domain with a pressure waveform
    std::vector<double> ap_sq;
    std::vector<Span> spans;
    bool has_orbitrap = false;

    Domains(const icarion_domain_desc* doms, const double* end_aperture_m, uint32_t count) : d(doms), n(count) {
        const double eps = 1e-12;   // DOMAIN_BOUNDARY_EPSILON (DomainManager.h:42)
        for (uint32_t i = 0; i < n; ++i) {
            const double ap = (end_aperture_m && end_aperture_m[i] > 0.0) ? end_aperture_m[i] : d[i].radius_m;
            ap_sq.push_back(ap * ap);
            if (d[i].instrument == ICARION_INSTRUMENT_ORBITRAP) {
                has_orbitrap = true;
                continue;
            }
            const double z0 = d[i].origin_m[2], z1 = z0 + d[i].length_m;
            spans.push_back(Span{std::min(z0, z1) - eps, std::max(z0, z1) + eps, d[i].radius_m + eps,
                                 static_cast<int>(i)});
        }
        std::sort(spans.begin(), spans.end(), [](const Span& a, const Span& b) { return a.z_min < b.z_min; });
    }
    int linear(V3 p) const {
        for (uint32_t i = 0; i < n; ++i)
            if (contains(d[i], p)) return static_cast<int>(i);
        return -1;
    }
    // find_domain_index (DomainManager.cpp:66-109)
    int find(V3 p) const {
        if (has_orbitrap || spans.empty()) return linear(p);
        auto it = std::lower_bound(spans.begin(), spans.end(), p.z,
                                   [](const Span& s, double z) { return s.z_max < z; });
        for (; it != spans.end(); ++it) {
            if (p.z < it->z_min) break;
            if ((p.x * p.x + p.y * p.y) > it->radius * it->radius) continue;
            const int idx = it->idx;
            if (idx >= 0 && idx < static_cast<int>(n) && contains(d[idx], p)) return idx;
        }
        return linear(p);
    }
    // forward_axial_bridge_domain (DomainManager.cpp:163-174)
    int bridge(int cur, V3 p) const {
        if (cur < 0 || static_cast<uint32_t>(cur + 1) >= n) return -1;
        const double cur_end = d[cur].origin_m[2] + d[cur].length_m;
        return (p.z >= cur_end) ? cur + 1 : -1;
    }
    // shared_boundary_handoff_domain (DomainManager.cpp:176-198), tolerance 1e-9 m
    int handoff(int cur, int found, V3 p) const {
        const double tol = 1e-9;
        if (cur < 0 || found != cur || static_cast<uint32_t>(cur + 1) >= n) return -1;
        const double local_z = p.z - d[cur].origin_m[2];
        if (local_z < d[cur].length_m - tol) return -1;
        const double dx = p.x - d[cur].origin_m[0];
        const double dy = p.y - d[cur].origin_m[1];
        return (dx * dx + dy * dy <= ap_sq[cur]) ? cur + 1 : -1;
    }
};

// One ion through the engine loop over several domains (SimulationEngine.cpp:
// 374-428; prepare / finalize of SimulationEngineLifecycle.cpp:72-195): the
// domain found at the step start carries the push and the collision; after
// the push the lookup, the bridge and the handoff decide life and the index.
uint64_t run_ion_multi(const Domains& D, const icarion_physics_desc& ph, const icarion_species_desc& sp,
                       uint32_t sidx, const double* atoms, IonCtx& ion, int& dom_idx, const double* t_table, double dt, uint64_t n_steps,
                       uint32_t ion_index, uint64_t step0, uint64_t& collisions, std::vector<V3>& scratch) {
    uint64_t pushed = 0;
    for (uint64_t k = 0; k < n_steps; ++k) {
        const double t = t_table[k];
        if (!ion.born && t >= ion.birth) {
            ion.born = 1;
            ion.active = 1;
        }
        if (!ion.active || !ion.born) continue;
        const int dom = D.find(ion.pos);
        if (dom < 0) {
            ion.active = 0;
            ion.death = t;
            continue;
        }
        dom_idx = dom;
        // the domain's environment of this step (with a pressure waveform: its density)
        const EnvSteps* env = D.env.empty() ? nullptr : D.env[dom].get();
        const icarion_domain_desc& d = env ? env->dom[k] : D.d[dom];
        const oracle_field* f = D.fields ? D.fields[dom] : nullptr;
        // the species row resolved for the step domain's gas: the handlers
        // and DampingForce take the ion's species with the environment of its
        // domain (SimulationEngineCollisionRuntime.cpp:36-46,
        // HSSCollisionHandler.cpp:263-310, EHSSCollisionHandler.cpp:462-547)
        const icarion_species_desc& sp_d = (D.species && D.species[dom]) ? D.species[dom][sidx] : sp;
        const double z_before = ion.pos.z;
        if (ph.integrator == ICARION_INTEGRATOR_BORIS)
            boris_step(d, f, ph.collision_model, sp_d, ion.m, ion.q, ion.pos, ion.vel, t, dt);
        else
            rk4_step(d, f, ph.collision_model, sp_d, ion.m, ion.q, ion.pos, ion.vel, t, dt);
        ++pushed;
        if (ph.collision_model != ICARION_COLLISION_NONE) {
            // the step domain's gas mixture, if it has one
            IonTables tb;
            const oracle_gas_mixture* gm = D.mixtures ? D.mixtures[dom] : nullptr;
            if (env && gm && gm->n_components > 0) gm = &env->mix[k];
            if (gm && gm->n_components > 0) {
                tb.mix = gm;
                tb.mix_row = gm->sigma_m2 + static_cast<size_t>(sidx) * gm->n_components;
            }
            PhiloxStream rng(ph.rng_seed, ion_index, step0 + k);
            const int ns = collision_substeps(ph);
            const double collision_dt = dt / static_cast<double>(ns);
            // the step domain's gas velocity at the pushed position
            // (SimulationEngineCollisionRuntime.cpp:93-101, 122-130)
            icarion_domain_desc dflow;
            const bool flow = f && f->flow_model != ICARION_FLOW_CONSTANT;
            if (flow) {
                dflow = d;
                const V3 u = gas_velocity_at(d, f, ion.pos);
                dflow.gas_velocity_m_s[0] = u.x;
                dflow.gas_velocity_m_s[1] = u.y;
                dflow.gas_velocity_m_s[2] = u.z;
            }
            const icarion_domain_desc& dc = flow ? dflow : d;
            for (int sub = 0; sub < ns; ++sub) {
                if (collide(dc, ph, sp_d, atoms, ion.m, ion.q, ion.vel, collision_dt, rng, scratch, tb)) {
                    ++ion.ncoll;
                    ++collisions;
                }
            }
        }
        if (std::isfinite(d.detector_z_m) && std::isnan(ion.arrival)) {
            if (z_before < d.detector_z_m && ion.pos.z >= d.detector_z_m)
                ion.arrival = t + dt * (d.detector_z_m - z_before) / (ion.pos.z - z_before);
        }
        const int found = D.find(ion.pos);
        if (found < 0) {
            const int b = D.bridge(dom_idx, ion.pos);
            if (b < 0) {
                ion.active = 0;
                ion.death = t;
                continue;
            }
            dom_idx = b;
        } else {
            const int h = D.handoff(dom_idx, found, ion.pos);
            dom_idx = h >= 0 ? h : found;
        }
        if (!finite3(ion.pos) || !finite3(ion.vel)) {
            ion.active = 0;
            ion.death = t;
        }
    }
    return pushed;
}

// Per-species views of a run's collision tables.
struct RunTables {
    const oracle_collision_tables* t = nullptr;
    std::vector<size_t> off_base, mu_base, ori_base;
    bool ok(uint32_t n_species) const {
        if (!t) return true;
        if (t->mixture && t->mixture->n_components > 0 && t->mixture->n_species != n_species) return false;
        if (t->ehss_offline && t->ehss_offline->n_species != n_species) return false;
        if (t->ehss_orient && t->ehss_orient->n_species != n_species) return false;
        return true;
    }
    explicit RunTables(const oracle_collision_tables* tables) : t(tables) {
        if (t && t->ehss_orient) {
            const oracle_ehss_orientations& o = *t->ehss_orient;
            ori_base.resize(o.n_species);
            size_t a = 0;
            for (uint32_t s = 0; s < o.n_species; ++s) {
                ori_base[s] = a;
                a += o.n_orient[s];
            }
        }
        if (t && t->ehss_offline) {
            const oracle_ehss_offline& o = *t->ehss_offline;
            off_base.resize(o.n_species);
            mu_base.resize(o.n_species);
            size_t a = 0, b = 0;
            for (uint32_t s = 0; s < o.n_species; ++s) {
                off_base[s] = a;
                mu_base[s] = b;
                a += o.n_orient[s];
                b += static_cast<size_t>(o.n_orient[s]) * o.mu_stride[s];
            }
        }
    }
    IonTables for_species(uint32_t s) const {
        IonTables tb;
        if (!t) return tb;
        if (t->mixture && t->mixture->n_components > 0) {
            tb.mix = t->mixture;
            tb.mix_row = t->mixture->sigma_m2 + static_cast<size_t>(s) * t->mixture->n_components;
        }
        if (t->ehss_offline && s < t->ehss_offline->n_species && t->ehss_offline->n_orient[s] > 0) {
            const oracle_ehss_offline& o = *t->ehss_offline;
            tb.off_n = o.n_orient[s];
            tb.off_stride = o.mu_stride[s];
            tb.off_sigma = o.sigma_eff + off_base[s];
            tb.off_mu = o.mu + mu_base[s];
        }
        if (t->ehss_orient && s < t->ehss_orient->n_species && t->ehss_orient->n_orient[s] > 0) {
            const oracle_ehss_orientations& o = *t->ehss_orient;
            tb.ori_n = o.n_orient[s];
            tb.ori_quat = o.quat_wxyz + 4 * ori_base[s];
            tb.ori_area = o.area_m2 + ori_base[s];
        }
        return tb;
    }
};

}  // namespace

extern "C" {

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key, out);
}

double oracle_uniform_from_words(uint32_t hi, uint32_t lo) { return uniform_from_words(hi, lo); }

void oracle_philox_uniforms(uint64_t seed, uint32_t ion, uint64_t step, uint32_t n, double* out) {
    PhiloxStream s(seed, ion, step);
    for (uint32_t i = 0; i < n; ++i) out[i] = s.uniform01();
}

void oracle_field_E(const icarion_domain_desc* domain, const oracle_field* field,
                    const double pos[3], double t, double E_out[3]) {
    const V3 E = field_E(*domain, field, V3{pos[0], pos[1], pos[2]}, t);
    E_out[0] = E.x; E_out[1] = E.y; E_out[2] = E.z;
}

int oracle_contains(const icarion_domain_desc* domain, const double pos[3]) {
    return contains(*domain, V3{pos[0], pos[1], pos[2]}) ? 1 : 0;
}

double oracle_orbitrap_r_for_z(double z, double R, double R_m) { return orbitrap_r_for_z(z, R, R_m); }

double oracle_waveform_value(const icarion_waveform_desc* w, double t) { return waveform_value(*w, t); }

void oracle_gas_velocity(const icarion_domain_desc* domain, const oracle_field* field, const double pos[3],
                         double out[3]) {
    const V3 u = gas_velocity_at(*domain, field, V3{pos[0], pos[1], pos[2]});
    out[0] = u.x;
    out[1] = u.y;
    out[2] = u.z;
}

double oracle_field_term_scale(const icarion_domain_desc* domain, const oracle_field_term* term,
                               const oracle_field* field, double t) {
    return term_scale(*domain, *term, t, field);
}

void oracle_interpolate_field(const oracle_field* field, const double pos[3], double E_out[3]) {
    const V3 E = interpolate(*field, V3{pos[0], pos[1], pos[2]});
    E_out[0] = E.x; E_out[1] = E.y; E_out[2] = E.z;
}

double oracle_damping_gamma(const icarion_domain_desc* domain, int collision_model,
                            const icarion_species_desc* sp, double mass_kg, double charge_C,
                            const double vel[3]) {
    return damping_gamma(*domain, collision_model, *sp, mass_kg, charge_C, V3{vel[0], vel[1], vel[2]});
}

void oracle_mt_sample_neutral_velocity(uint64_t seed, double T_K, double mass_kg,
                                       const double flow[3], double out[3]) {
    MtRng rng(seed);
    const V3 v = sample_neutral_velocity(T_K, mass_kg, V3{flow[0], flow[1], flow[2]}, rng);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

void oracle_mt_hss_collision(uint64_t seed, const double v_ion[3], const double v_neutral[3],
                             double ion_mass, double neutral_mass, double out[3]) {
    MtRng rng(seed);
    const V3 v = hss_collision(V3{v_ion[0], v_ion[1], v_ion[2]},
                               V3{v_neutral[0], v_neutral[1], v_neutral[2]}, ion_mass,
                               neutral_mass, rng);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

int oracle_mt_hss_mixture(uint64_t seed, const icarion_domain_desc* domain,
                          const oracle_gas_mixture* mixture, const double* sigma_row, double mass_kg,
                          double vel[3], double dt, double vn_out[3]) {
    MtRng rng(seed);
    V3 v{vel[0], vel[1], vel[2]};
    V3 vn{0.0, 0.0, 0.0};
    const int sel = collide_mixture(*domain, *mixture, sigma_row, mass_kg, v, dt, rng, &vn);
    vel[0] = v.x; vel[1] = v.y; vel[2] = v.z;
    if (vn_out) {
        vn_out[0] = vn.x; vn_out[1] = vn.y; vn_out[2] = vn.z;
    }
    return sel;
}

int oracle_mt_ehss_offline(uint64_t seed, const icarion_domain_desc* domain, uint32_t n_orient,
                           uint32_t mu_stride, const double* sigma_eff, const double* mu,
                           double mass_kg, double vel[3], double dt, uint32_t* orient_out,
                           double* mu_out, double vn_out[3]) {
    MtRng rng(seed);
    V3 v{vel[0], vel[1], vel[2]};
    V3 vn{0.0, 0.0, 0.0};
    uint32_t o = 0;
    double m = NAN;
    const bool hit = ehss_offline_collide(*domain, n_orient, mu_stride, sigma_eff, mu, mass_kg, v, dt,
                                          rng, &o, &m, &vn);
    vel[0] = v.x; vel[1] = v.y; vel[2] = v.z;
    if (orient_out) *orient_out = o;
    if (mu_out) *mu_out = m;
    if (vn_out) {
        vn_out[0] = vn.x; vn_out[1] = vn.y; vn_out[2] = vn.z;
    }
    return hit ? 1 : 0;
}

int oracle_mt_ehss_mixture(uint64_t seed, const icarion_domain_desc* domain, const oracle_gas_mixture* mix,
                           uint32_t species, const double* atoms_xyzr, uint32_t n_atoms, double mass_kg,
                           double vel[3], double dt) {
    MtRng rng(seed);
    std::vector<V3> scratch;
    V3 v{vel[0], vel[1], vel[2]};
    const double* row = mix->sigma_m2 + static_cast<size_t>(species) * mix->n_components;
    const bool hit = ehss_mixture_collide(*domain, *mix, row, atoms_xyzr, n_atoms, mass_kg, v, dt, rng, scratch);
    vel[0] = v.x; vel[1] = v.y; vel[2] = v.z;
    return hit ? 1 : 0;
}

void oracle_mt_ehss_oriented(uint64_t seed, const double v_ion[3], const double v_neutral[3], double ion_mass,
                             double neutral_mass, double ion_radius, const double* atoms_xyzr, uint32_t n_atoms,
                             const double quat_wxyz[4], double sigma_eff, double out[3]) {
    MtRng rng(seed);
    std::vector<V3> scratch;
    double R[3][3];
    quaternion_to_rotation(quat_wxyz, R);
    const V3 v = ehss_collision_oriented(V3{v_ion[0], v_ion[1], v_ion[2]},
                                         V3{v_neutral[0], v_neutral[1], v_neutral[2]}, ion_mass, neutral_mass,
                                         ion_radius, atoms_xyzr, n_atoms, R, sigma_eff, kEhssMaxAttempts * 16, rng,
                                         scratch);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

void oracle_quaternion_to_rotation(const double quat_wxyz[4], double R[9]) {
    double M[3][3];
    quaternion_to_rotation(quat_wxyz, M);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[3 * i + j] = M[i][j];
}

void oracle_mt_ehss_collision(uint64_t seed, const double v_ion[3], const double v_neutral[3],
                              double ion_mass, double neutral_mass, double ion_radius,
                              const double* atoms_xyzr, uint32_t n_atoms, double out[3]) {
    MtRng rng(seed);
    std::vector<V3> scratch;
    const V3 v = ehss_collision(V3{v_ion[0], v_ion[1], v_ion[2]},
                                V3{v_neutral[0], v_neutral[1], v_neutral[2]}, ion_mass,
                                neutral_mass, ion_radius, atoms_xyzr, n_atoms, rng, scratch);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

double oracle_oapa_sigma(const double* atoms, uint32_t n_atoms, double neutral_radius) {
    MtRng rng(12345);
    double A_sum = 0.0;
    double R[3][3];
    const int n_orientations = 64;
    for (int k = 0; k < n_orientations; ++k) {
        random_rotation(rng, R);
        double Rmax = 0.0;
        for (uint32_t i = 0; i < n_atoms; ++i) {
            const double cx = atoms[4 * i], cy = atoms[4 * i + 1], cz = atoms[4 * i + 2];
            const V3 c{R[0][0] * cx + R[0][1] * cy + R[0][2] * cz,
                       R[1][0] * cx + R[1][1] * cy + R[1][2] * cz,
                       R[2][0] * cx + R[2][1] * cy + R[2][2] * cz};
            const double r_eff = std::sqrt(c.x * c.x + c.y * c.y) + atoms[4 * i + 3] + neutral_radius;
            Rmax = std::max(Rmax, r_eff);
        }
        A_sum += kPi * Rmax * Rmax;
    }
    return A_sum / static_cast<double>(n_orientations);
}

void oracle_push_one(const icarion_domain_desc* domain, const icarion_physics_desc* physics,
                     const icarion_species_desc* sp, const oracle_field* field, double mass_kg,
                     double charge_C, double pos[3], double vel[3], double t, double dt) {
    V3 p{pos[0], pos[1], pos[2]}, v{vel[0], vel[1], vel[2]};
    if (physics->integrator == ICARION_INTEGRATOR_BORIS)
        boris_step(*domain, field, physics->collision_model, *sp, mass_kg, charge_C, p, v, t, dt);
    else
        rk4_step(*domain, field, physics->collision_model, *sp, mass_kg, charge_C, p, v, t, dt);
    pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
    vel[0] = v.x; vel[1] = v.y; vel[2] = v.z;
}

int oracle_find_domain(const icarion_domain_desc* domains, const double* end_aperture_m, uint32_t n_domains,
                       const double pos[3]) {
    const Domains D(domains, end_aperture_m, n_domains);
    return D.find(V3{pos[0], pos[1], pos[2]});
}

int oracle_run_multi(const icarion_domain_desc* domains, const double* end_aperture_m, uint32_t n_domains,
                     const oracle_field* const* fields, const oracle_gas_mixture* const* mixtures,
                     const icarion_species_desc* const* domain_species,
                     const icarion_physics_desc* physics,
                     const icarion_species_desc* species, uint32_t n_species, const double* atoms,
                     icarion_ion_soa* ions, int32_t* domain_index, double t0, double dt, uint64_t n_steps,
                     uint64_t step_offset, int n_threads, icarion_step_stats* stats) {
    if (!domains || n_domains == 0 || !physics || !species || !ions || !domain_index || n_species == 0)
        return ICARION_ERR_INVALID;
    Domains D(domains, end_aperture_m, n_domains);
    D.fields = fields;
    D.mixtures = mixtures;
    D.species = domain_species;
    const size_t n = ions->n;
    std::vector<double> t_table(n_steps + 1);
    double t = t0;
    for (uint64_t k = 0; k <= n_steps; ++k) {
        t_table[k] = t;
        t = t + dt;
    }
    for (uint32_t k = 0; fields && k < n_domains; ++k) {
        if (!fields[k] || !fields[k]->pressure) continue;
        D.env.resize(n_domains);
        D.env[k] = std::make_unique<EnvSteps>(domains[k], *fields[k]->pressure, mixtures ? mixtures[k] : nullptr,
                                              t_table.data(), n_steps);
        if (!D.env[k]->ok) return ICARION_ERR_INVALID;
    }
    uint64_t pushed_total = 0, coll_total = 0, alive = 0, arrived = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    const long long chunk = omp_chunk(n);
    #pragma omp parallel reduction(+ : pushed_total, coll_total, alive, arrived)
    {
        std::vector<V3> scratch;
        #pragma omp for schedule(dynamic, chunk)
        for (long long ii = 0; ii < static_cast<long long>(n); ++ii) {
            const size_t i = static_cast<size_t>(ii);
            const uint32_t s = ions->species[i];
            const icarion_species_desc& sp = species[s < n_species ? s : 0];
            IonCtx ion;
            ion.pos = V3{ions->x[i], ions->y[i], ions->z[i]};
            ion.vel = V3{ions->vx[i], ions->vy[i], ions->vz[i]};
            ion.m = ions->mass_kg[i];
            ion.q = ions->charge_C[i];
            ion.active = ions->active[i];
            ion.born = ions->born ? ions->born[i] : 1;
            ion.birth = ions->birth_time_s ? ions->birth_time_s[i] : t0;
            ion.death = ions->death_time_s ? ions->death_time_s[i] : -1.0;
            ion.arrival = ions->arrival_time_s ? ions->arrival_time_s[i] : NAN;
            ion.ncoll = ions->n_collisions ? ions->n_collisions[i] : 0;
            int dom = domain_index[i];
            uint64_t coll = 0;
            const uint64_t pushed = run_ion_multi(D, *physics, sp, s < n_species ? s : 0, atoms, ion, dom, t_table.data(), dt, n_steps,
                                                  static_cast<uint32_t>(i), step_offset, coll, scratch);
            ions->x[i] = ion.pos.x; ions->y[i] = ion.pos.y; ions->z[i] = ion.pos.z;
            ions->vx[i] = ion.vel.x; ions->vy[i] = ion.vel.y; ions->vz[i] = ion.vel.z;
            ions->active[i] = ion.active;
            if (ions->born) ions->born[i] = ion.born;
            if (ions->death_time_s) ions->death_time_s[i] = ion.death;
            if (ions->arrival_time_s) ions->arrival_time_s[i] = ion.arrival;
            if (ions->n_collisions) ions->n_collisions[i] = ion.ncoll;
            domain_index[i] = dom;
            pushed_total += pushed;
            coll_total += coll;
            if (ion.active && ion.born) ++alive;
            if (!std::isnan(ion.arrival)) ++arrived;
        }
    }
    if (stats) {
        stats->ion_steps = pushed_total;
        stats->collisions = coll_total;
        stats->alive = alive;
        stats->arrived = arrived;
        stats->t_end_s = t_table[n_steps];
        stats->rejected = 0;
        stats->device_ms = 0.0;
    }
    return ICARION_OK;
}

int oracle_run(const icarion_domain_desc* domain, const icarion_physics_desc* physics,
               const icarion_species_desc* species, uint32_t n_species, const double* atoms,
               uint32_t n_atoms, const oracle_field* field, icarion_ion_soa* ions, double t0,
               double dt, uint64_t n_steps, uint64_t step_offset, int rng_kind, int n_threads,
               icarion_step_stats* stats, const oracle_collision_tables* tables) {
    (void)n_atoms;
    if (!domain || !physics || !species || !ions || n_species == 0) return ICARION_ERR_INVALID;
    const RunTables rt(tables);
    if (!rt.ok(n_species)) return ICARION_ERR_INVALID;
    const size_t n = ions->n;
    // Engine time t_k: the reference advances every live ion by t_i += dt and
    // takes the engine time as max_i t_i (SimulationEngineLifecycle.cpp:172,
    // SimulationEngineUtils.cpp:49-62), i.e. the running sum t0 + dt + dt + ...
    std::vector<double> t_table(n_steps + 1);
    double t = t0;
    for (uint64_t k = 0; k <= n_steps; ++k) {
        t_table[k] = t;
        t = t + dt;
    }
    std::unique_ptr<EnvSteps> env;
    if (field && field->pressure) {
        env = std::make_unique<EnvSteps>(*domain, *field->pressure, tables ? tables->mixture : nullptr,
                                         t_table.data(), n_steps);
        if (!env->ok) return ICARION_ERR_INVALID;
    }
    uint64_t pushed_total = 0, coll_total = 0, alive = 0, arrived = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    const long long chunk = omp_chunk(n);
    #pragma omp parallel reduction(+ : pushed_total, coll_total, alive, arrived)
    {
        std::vector<V3> scratch;
        #pragma omp for schedule(dynamic, chunk)
        for (long long ii = 0; ii < static_cast<long long>(n); ++ii) {
            const size_t i = static_cast<size_t>(ii);
            const uint32_t s = ions->species[i];
            const icarion_species_desc& sp = species[s < n_species ? s : 0];
            IonCtx ion;
            ion.pos = V3{ions->x[i], ions->y[i], ions->z[i]};
            ion.vel = V3{ions->vx[i], ions->vy[i], ions->vz[i]};
            ion.m = ions->mass_kg[i];
            ion.q = ions->charge_C[i];
            ion.active = ions->active[i];
            ion.born = ions->born ? ions->born[i] : 1;
            ion.birth = ions->birth_time_s ? ions->birth_time_s[i] : t0;
            ion.death = ions->death_time_s ? ions->death_time_s[i] : -1.0;
            ion.arrival = ions->arrival_time_s ? ions->arrival_time_s[i] : NAN;
            ion.ncoll = ions->n_collisions ? ions->n_collisions[i] : 0;
            uint64_t coll = 0;
            uint64_t pushed = 0;
            const IonTables tb = rt.for_species(s < n_species ? s : 0);
            if (rng_kind == ORACLE_RNG_MT19937_64) {
                MtRng rng(physics->rng_seed + static_cast<uint64_t>(i));
                pushed = run_ion(*domain, *physics, sp, atoms, field, ion, t_table.data(), dt,
                                 n_steps, static_cast<uint32_t>(i), step_offset, &rng, coll,
                                 scratch, tb, env.get());
            } else {
                pushed = run_ion(*domain, *physics, sp, atoms, field, ion, t_table.data(), dt,
                                 n_steps, static_cast<uint32_t>(i), step_offset, nullptr, coll,
                                 scratch, tb, env.get());
            }
            ions->x[i] = ion.pos.x; ions->y[i] = ion.pos.y; ions->z[i] = ion.pos.z;
            ions->vx[i] = ion.vel.x; ions->vy[i] = ion.vel.y; ions->vz[i] = ion.vel.z;
            ions->active[i] = ion.active;
            if (ions->born) ions->born[i] = ion.born;
            if (ions->death_time_s) ions->death_time_s[i] = ion.death;
            if (ions->arrival_time_s) ions->arrival_time_s[i] = ion.arrival;
            if (ions->n_collisions) ions->n_collisions[i] = ion.ncoll;
            pushed_total += pushed;
            coll_total += coll;
            if (ion.active && ion.born) ++alive;
            if (!std::isnan(ion.arrival)) ++arrived;
        }
    }
    if (stats) {
        stats->ion_steps = pushed_total;
        stats->collisions = coll_total;
        stats->alive = alive;
        stats->arrived = arrived;
        stats->t_end_s = t_table[n_steps];
        stats->device_ms = 0.0;
    }
    return ICARION_OK;
}

double oracle_rk45_error(const icarion_rk45_desc* cfg, const double y4[6], const double y5[6],
                         const double y0[6]) {
    return rk45_error(*cfg, y4, y5, y0);
}

double oracle_rk45_new_step(const icarion_rk45_desc* cfg, double dt, double err, double dt_min,
                            double dt_max, double prev_err) {
    return rk45_new_step(*cfg, dt, err, dt_min, dt_max, prev_err);
}

double oracle_rk45_push_one(const icarion_domain_desc* domain, const icarion_physics_desc* physics,
                            const icarion_rk45_desc* cfg, const icarion_species_desc* sp,
                            const oracle_field* field, double mass_kg, double charge_C,
                            double pos[3], double vel[3], double t, double dt,
                            oracle_rk45_ion* st, double* dt_next, int* rejected) {
    V3 p{pos[0], pos[1], pos[2]}, v{vel[0], vel[1], vel[2]};
    double dn = 0.0;
    int rej = 0;
    const double used = rk45_step(*domain, field, physics->collision_model, *cfg, *sp, mass_kg,
                                  charge_C, p, v, t, dt, *st, dn, rej);
    pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
    vel[0] = v.x; vel[1] = v.y; vel[2] = v.z;
    if (dt_next) *dt_next = dn;
    if (rejected) *rejected = rej;
    return used;
}

int oracle_run_rk45(const icarion_domain_desc* domain, const icarion_physics_desc* physics,
                    const icarion_rk45_desc* cfg, const icarion_species_desc* species,
                    uint32_t n_species, const double* atoms, uint32_t n_atoms,
                    const oracle_field* field, icarion_ion_soa* ions, oracle_rk45_ion* rk45,
                    double t0, double dt, uint64_t n_steps, uint64_t step_offset, int n_threads,
                    icarion_step_stats* stats, const oracle_collision_tables* tables) {
    (void)n_atoms;
    if (!domain || !physics || !cfg || !species || !ions || !rk45 || n_species == 0)
        return ICARION_ERR_INVALID;
    const RunTables rt(tables);
    if (!rt.ok(n_species)) return ICARION_ERR_INVALID;
    const icarion_domain_desc& d = *domain;
    const icarion_physics_desc& ph = *physics;
    const size_t n = ions->n;
    const bool detect = std::isfinite(d.detector_z_m);
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    for (size_t i = 0; i < n; ++i) {
        if (std::isnan(rk45[i].t)) rk45[i].t = t0;
        if (!(rk45[i].dt > 0.0)) rk45[i].dt = dt;
    }
    double T = t0;
    uint64_t pushed_total = 0, coll_total = 0, rejected_total = 0;
    bool failed = false;
    const long long chunk = omp_chunk(n);
    for (uint64_t k = 0; k < n_steps; ++k) {
        double max_dt_used = 0.0;
        // the step's environment: with a pressure waveform, its density at the
        // engine time T (update_dynamic_environments)
        std::unique_ptr<EnvSteps> env;
        if (field && field->pressure) {
            env = std::make_unique<EnvSteps>(d, *field->pressure, tables ? tables->mixture : nullptr, &T, 1);
            if (!env->ok) return ICARION_ERR_INVALID;
        }
        const icarion_domain_desc& de = env ? env->dom[0] : d;
        const bool flow = field && field->flow_model != ICARION_FLOW_CONSTANT;
        #pragma omp parallel reduction(+ : pushed_total, coll_total, rejected_total) \
            reduction(max : max_dt_used) reduction(|| : failed)
        {
            std::vector<V3> scratch;
            #pragma omp for schedule(dynamic, chunk)
            for (long long ii = 0; ii < static_cast<long long>(n); ++ii) {
                const size_t i = static_cast<size_t>(ii);
                const uint32_t s = ions->species[i];
                const icarion_species_desc& sp = species[s < n_species ? s : 0];
                oracle_rk45_ion& st = rk45[i];
                uint8_t born = ions->born ? ions->born[i] : 1;
                uint8_t active = ions->active[i];
                const double birth = ions->birth_time_s ? ions->birth_time_s[i] : t0;
                // prepare_ions_for_integration (SimulationEngineLifecycle.cpp:86-107)
                if (!born && T >= birth) {
                    born = 1;
                    active = 1;
                }
                if (ions->born) ions->born[i] = born;
                if (!active || !born) {
                    ions->active[i] = active;
                    continue;
                }
                V3 pos{ions->x[i], ions->y[i], ions->z[i]};
                V3 vel{ions->vx[i], ions->vy[i], ions->vz[i]};
                if (!contains(d, pos)) {
                    ions->active[i] = 0;
                    if (ions->death_time_s) ions->death_time_s[i] = T;
                    continue;
                }
                const double z_before = pos.z;
                double dt_next = 0.0;
                int rej = 0;
                const double used = rk45_step(de, field, ph.collision_model, *cfg, sp,
                                              ions->mass_kg[i], ions->charge_C[i], pos, vel, T,
                                              st.dt, st, dt_next, rej);
                rejected_total += static_cast<uint64_t>(rej);
                if (used < 0.0) {
                    failed = true;
                    continue;
                }
                ++pushed_total;
                max_dt_used = std::max(max_dt_used, used);
                // perform_collisions with the accepted dt
                if (ph.collision_model != ICARION_COLLISION_NONE) {
                    PhiloxStream rng(ph.rng_seed, static_cast<uint32_t>(i), step_offset + k);
                    IonTables tb = rt.for_species(s < n_species ? s : 0);
                    if (env && tb.mix) tb.mix = &env->mix[0];
                    // the gas velocity at the pushed position
                    icarion_domain_desc dflow;
                    if (flow) {
                        dflow = de;
                        const V3 u = gas_velocity_at(de, field, pos);
                        dflow.gas_velocity_m_s[0] = u.x;
                        dflow.gas_velocity_m_s[1] = u.y;
                        dflow.gas_velocity_m_s[2] = u.z;
                    }
                    const icarion_domain_desc& dc = flow ? dflow : de;
                    // substeps of the ion's own dt (dt_used_per_ion / substeps)
                    const int ns = collision_substeps(ph);
                    const double collision_dt = used / static_cast<double>(ns);
                    for (int sub = 0; sub < ns; ++sub) {
                        if (collide(dc, ph, sp, atoms, ions->mass_kg[i], ions->charge_C[i], vel, collision_dt,
                                    rng, scratch, tb)) {
                            if (ions->n_collisions) ++ions->n_collisions[i];
                            ++coll_total;
                        }
                    }
                }
                if (detect && ions->arrival_time_s && std::isnan(ions->arrival_time_s[i])) {
                    if (z_before < d.detector_z_m && pos.z >= d.detector_z_m)
                        ions->arrival_time_s[i] =
                            T + used * (d.detector_z_m - z_before) / (pos.z - z_before);
                }
                // finalize_ions_after_integration (SimulationEngineLifecycle.cpp:137-192)
                bool alive = contains(d, pos);
                if (alive) {
                    st.t = st.t + used;
                    st.dt = dt_next > 0.0 ? dt_next : (used > 0.0 ? used : dt);
                    if (!finite3(pos) || !finite3(vel)) alive = false;
                }
                if (!alive) {
                    active = 0;
                    if (ions->death_time_s) ions->death_time_s[i] = T;
                }
                ions->active[i] = active;
                ions->x[i] = pos.x; ions->y[i] = pos.y; ions->z[i] = pos.z;
                ions->vx[i] = vel.x; ions->vy[i] = vel.y; ions->vz[i] = vel.z;
            }
        }
        if (failed) return ICARION_ERR_INVALID;
        // next_engine_time_after_step (SimulationEngineUtils.cpp:49-83), with
        // total_time_s unbounded
        double Tn = T;
        for (size_t i = 0; i < n; ++i) Tn = std::max(Tn, rk45[i].t);
        if (!(Tn > T)) {
            double next_birth = INFINITY;
            for (size_t i = 0; i < n; ++i) {
                if (ions->born && ions->born[i]) continue;
                const double bt = ions->birth_time_s ? ions->birth_time_s[i] : t0;
                if (bt > T) next_birth = std::min(next_birth, bt);
            }
            if (std::isfinite(next_birth)) Tn = next_birth;
            else if (max_dt_used > 0.0) Tn = T + max_dt_used;
            else Tn = T + dt;
        }
        T = Tn;
    }
    if (ions->time_s)
        for (size_t i = 0; i < n; ++i) ions->time_s[i] = rk45[i].t;
    if (ions->dt_s)
        for (size_t i = 0; i < n; ++i) ions->dt_s[i] = rk45[i].dt;
    if (stats) {
        uint64_t alive = 0, arrived = 0;
        for (size_t i = 0; i < n; ++i) {
            if (ions->active[i] && (!ions->born || ions->born[i])) ++alive;
            if (ions->arrival_time_s && !std::isnan(ions->arrival_time_s[i])) ++arrived;
        }
        stats->ion_steps = pushed_total;
        stats->collisions = coll_total;
        stats->alive = alive;
        stats->arrived = arrived;
        stats->rejected = rejected_total;
        stats->t_end_s = T;
        stats->device_ms = 0.0;
    }
    return ICARION_OK;
}

}  // extern "C"

// ---- EHSS offline precompute and surrogate reference ------------------------
namespace {

constexpr int kH = 128;
constexpr int kIn = 6;
constexpr int kOut = 3;
constexpr size_t kOffW1 = 0, kOffB1 = kOffW1 + kH * kIn, kOffW2 = kOffB1 + kH,
                 kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kOut * kH,
                 kNParams = kOffB3 + kOut;
static_assert(kNParams == ORACLE_MLP_NPARAMS, "parameter count");

// quaternion_to_rotation + sample_random_quaternion (tools/ehss_offline_precompute.cpp:115-156)
template <class Rng>
void precompute_rotation(Rng& rng, double R[3][3]) {
    const double u1 = rng.uniform01();
    const double u2 = rng.uniform01();
    const double u3 = rng.uniform01();
    double x = std::sqrt(1.0 - u1) * std::sin(2.0 * kPi * u2);
    double y = std::sqrt(1.0 - u1) * std::cos(2.0 * kPi * u2);
    double z = std::sqrt(u1) * std::sin(2.0 * kPi * u3);
    double w = std::sqrt(u1) * std::cos(2.0 * kPi * u3);
    const double nrm = std::sqrt(w * w + x * x + y * y + z * z);
    w /= nrm; x /= nrm; y /= nrm; z /= nrm;
    R[0][0] = 1.0 - 2.0 * (y * y + z * z);
    R[0][1] = 2.0 * (x * y - z * w);
    R[0][2] = 2.0 * (x * z + y * w);
    R[1][0] = 2.0 * (x * y + z * w);
    R[1][1] = 1.0 - 2.0 * (x * x + z * z);
    R[1][2] = 2.0 * (y * z - x * w);
    R[2][0] = 2.0 * (x * z - y * w);
    R[2][1] = 2.0 * (y * z + x * w);
    R[2][2] = 1.0 - 2.0 * (x * x + y * y);
}

// detect_atom_collision with collision axis +z (tools/ehss_offline_precompute.cpp:162-213)
bool precompute_contact(const std::vector<V3>& ra, const std::vector<double>& rad, double ox,
                        double oy, double r_gas, double& mu) {
    bool hit = false;
    double best_s = 0.0;
    double nz_best = 0.0;
    for (size_t j = 0; j < ra.size(); ++j) {
        const double Rsum = rad[j] + r_gas;
        const double Rsum2 = Rsum * Rsum;
        const double rx = ox - ra[j].x, ry = oy - ra[j].y, rz = 0.0 - ra[j].z;
        const double sstar = rz;
        const double dmin2 = rx * rx + ry * ry;
        if (dmin2 <= Rsum2) {
            const double h = std::sqrt(std::max(0.0, Rsum2 - dmin2));
            const double s_hit = sstar - h;
            const double pz = rz - s_hit;
            const double phr2 = rx * rx + ry * ry + pz * pz;
            double nz;
            if (phr2 <= kMinContactDistSq) {
                nz = -1.0;
            } else {
                const double inv = 1.0 / std::sqrt(phr2);
                nz = pz * inv;
                if (nz > 0.0) nz = -nz;
            }
            if (!hit || s_hit < best_s) {
                hit = true;
                best_s = s_hit;
                nz_best = nz;
            }
        }
    }
    if (hit) mu = std::min(1.0, std::max(-1.0, -nz_best));
    return hit;
}

}  // namespace

extern "C" {

void oracle_ehss_precompute(const float* atoms, const uint32_t* offsets, const uint32_t* counts,
                            uint32_t n_clusters, double r_gas, uint32_t n_orient,
                            uint32_t n_sigma, uint32_t n_mu, uint32_t max_tries, uint64_t seed,
                            double* sigma_out, double* mu_out) {
    const uint64_t key = seed ^ ORACLE_PRECOMPUTE_KEY;
    #pragma omp parallel for schedule(dynamic, 4) collapse(2)
    for (long long c = 0; c < static_cast<long long>(n_clusters); ++c) {
        for (long long k = 0; k < static_cast<long long>(n_orient); ++k) {
            PhiloxStream rng(key, static_cast<uint32_t>(c), static_cast<uint64_t>(k));
            double R[3][3];
            precompute_rotation(rng, R);
            const uint32_t a0 = offsets[c], na = counts[c];
            std::vector<V3> ra(na);
            std::vector<double> rad(na);
            double b_base = 0.0;
            for (uint32_t j = 0; j < na; ++j) {
                const double cx = atoms[4 * (a0 + j)], cy = atoms[4 * (a0 + j) + 1],
                             cz = atoms[4 * (a0 + j) + 2];
                rad[j] = atoms[4 * (a0 + j) + 3];
                ra[j] = V3{R[0][0] * cx + R[0][1] * cy + R[0][2] * cz,
                           R[1][0] * cx + R[1][1] * cy + R[1][2] * cz,
                           R[2][0] * cx + R[2][1] * cy + R[2][2] * cz};
                b_base = std::max(b_base, std::sqrt(ra[j].x * ra[j].x + ra[j].y * ra[j].y) + rad[j]);
            }
            const double b_max = b_base + r_gas;
            const double area = kPi * b_max * b_max;
            uint32_t hits = 0;
            double mu;
            for (uint32_t s = 0; s < n_sigma; ++s) {
                const double u1 = rng.uniform01();
                const double u2 = rng.uniform01();
                const double r = b_max * std::sqrt(u1);
                const double phi = 2.0 * kPi * u2;
                if (precompute_contact(ra, rad, r * std::cos(phi), r * std::sin(phi), r_gas, mu)) ++hits;
            }
            sigma_out[c * n_orient + k] =
                n_sigma ? (static_cast<double>(hits) / static_cast<double>(n_sigma)) * area : 0.0;
            uint32_t got = 0, tries = 0;
            double* row = mu_out + (static_cast<size_t>(c) * n_orient + k) * n_mu;
            while (got < n_mu && tries < max_tries) {
                ++tries;
                const double u1 = rng.uniform01();
                const double u2 = rng.uniform01();
                const double r = b_max * std::sqrt(u1);
                const double phi = 2.0 * kPi * u2;
                if (precompute_contact(ra, rad, r * std::cos(phi), r * std::sin(phi), r_gas, mu)) row[got++] = mu;
            }
            for (; got < n_mu; ++got) row[got] = NAN;
        }
    }
}

void oracle_oapa_sigma_f32(const float* atoms, const uint32_t* offsets, const uint32_t* counts,
                           uint32_t n_clusters, double r_gas, double* sigma_out) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (long long c = 0; c < static_cast<long long>(n_clusters); ++c) {
        std::vector<double> a(4 * static_cast<size_t>(counts[c]));
        for (size_t i = 0; i < a.size(); ++i) a[i] = atoms[4 * static_cast<size_t>(offsets[c]) + i];
        sigma_out[c] = oracle_oapa_sigma(a.data(), counts[c], r_gas);
    }
}

void oracle_ehss_features(const float* atoms, const uint32_t* offsets, const uint32_t* counts,
                          uint32_t n_clusters, double r_gas, const double* sigma_oapa,
                          float* f) {
    for (uint32_t c = 0; c < n_clusters; ++c) {
        const uint32_t a0 = offsets[c], na = counts[c];
        double mx = 0, my = 0, mz = 0, mr = 0;
        for (uint32_t j = 0; j < na; ++j) {
            mx += atoms[4 * (a0 + j)];
            my += atoms[4 * (a0 + j) + 1];
            mz += atoms[4 * (a0 + j) + 2];
            mr += atoms[4 * (a0 + j) + 3];
        }
        mx /= na; my /= na; mz /= na; mr /= na;
        double rg2 = 0;
        for (uint32_t j = 0; j < na; ++j) {
            const double dx = atoms[4 * (a0 + j)] - mx, dy = atoms[4 * (a0 + j) + 1] - my,
                         dz = atoms[4 * (a0 + j) + 2] - mz;
            rg2 += dx * dx + dy * dy + dz * dz;
        }
        const double b_eff = std::sqrt(sigma_oapa[c] / kPi);
        f[3 * c] = static_cast<float>(std::log(static_cast<double>(na)) / std::log(500.0));
        f[3 * c + 1] = static_cast<float>(std::sqrt(rg2 / na) / b_eff);
        f[3 * c + 2] = static_cast<float>((mr + r_gas) / b_eff);
    }
}

void oracle_mlp_forward(const float* p, const float* X, size_t n, float* Y) {
    #pragma omp parallel for schedule(static)
    for (long long s = 0; s < static_cast<long long>(n); ++s) {
        double h1[kH], h2[kH];
        for (int o = 0; o < kH; ++o) {
            double a = p[kOffB1 + o];
            for (int i = 0; i < kIn; ++i) a += static_cast<double>(p[kOffW1 + o * kIn + i]) * X[s * kIn + i];
            h1[o] = a > 0.0 ? a : 0.0;
        }
        for (int o = 0; o < kH; ++o) {
            double a = p[kOffB2 + o];
            for (int i = 0; i < kH; ++i) a += static_cast<double>(p[kOffW2 + o * kH + i]) * h1[i];
            h2[o] = a > 0.0 ? a : 0.0;
        }
        for (int o = 0; o < kOut; ++o) {
            double a = p[kOffB3 + o];
            for (int i = 0; i < kH; ++i) a += static_cast<double>(p[kOffW3 + o * kH + i]) * h2[i];
            Y[s * kOut + o] = static_cast<float>(a);
        }
    }
}

double oracle_mlp_loss_grad(const float* p, const float* X, const float* T, size_t n, double* g) {
    for (size_t i = 0; i < kNParams; ++i) g[i] = 0.0;
    double loss = 0.0;
    const double inv_n = 1.0 / static_cast<double>(n);
    for (size_t s = 0; s < n; ++s) {
        double z1[kH], h1[kH], z2[kH], h2[kH], y[kOut], dy[kOut], dh2[kH], dh1[kH];
        for (int o = 0; o < kH; ++o) {
            double a = p[kOffB1 + o];
            for (int i = 0; i < kIn; ++i) a += static_cast<double>(p[kOffW1 + o * kIn + i]) * X[s * kIn + i];
            z1[o] = a;
            h1[o] = a > 0.0 ? a : 0.0;
        }
        for (int o = 0; o < kH; ++o) {
            double a = p[kOffB2 + o];
            for (int i = 0; i < kH; ++i) a += static_cast<double>(p[kOffW2 + o * kH + i]) * h1[i];
            z2[o] = a;
            h2[o] = a > 0.0 ? a : 0.0;
        }
        for (int o = 0; o < kOut; ++o) {
            double a = p[kOffB3 + o];
            for (int i = 0; i < kH; ++i) a += static_cast<double>(p[kOffW3 + o * kH + i]) * h2[i];
            y[o] = a;
            const double d = a - T[s * kOut + o];
            loss += 0.5 * d * d * inv_n;
            dy[o] = d * inv_n;
        }
        for (int i = 0; i < kH; ++i) dh2[i] = 0.0;
        for (int o = 0; o < kOut; ++o) {
            g[kOffB3 + o] += dy[o];
            for (int i = 0; i < kH; ++i) {
                g[kOffW3 + o * kH + i] += dy[o] * h2[i];
                dh2[i] += dy[o] * p[kOffW3 + o * kH + i];
            }
        }
        for (int i = 0; i < kH; ++i) dh1[i] = 0.0;
        for (int o = 0; o < kH; ++o) {
            const double dz = z2[o] > 0.0 ? dh2[o] : 0.0;
            g[kOffB2 + o] += dz;
            for (int i = 0; i < kH; ++i) {
                g[kOffW2 + o * kH + i] += dz * h1[i];
                dh1[i] += dz * p[kOffW2 + o * kH + i];
            }
        }
        for (int o = 0; o < kH; ++o) {
            const double dz = z1[o] > 0.0 ? dh1[o] : 0.0;
            g[kOffB1 + o] += dz;
            for (int i = 0; i < kIn; ++i) g[kOffW1 + o * kIn + i] += dz * X[s * kIn + i];
        }
        (void)y;
    }
    return loss;
}

}  // extern "C"
