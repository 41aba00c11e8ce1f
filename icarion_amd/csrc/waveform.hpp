// waveform.hpp - evaluation of time-varying voltages (ValueOrWaveform:
// the waveform types of WaveformConfig.h:19-200 and the interpolated table of
// WaveformConfig.cpp:10-91).  The fixed-step paths evaluate them on the host
// at every stage time into the FP64 step table; RK45 evaluates them in the
// kernel at each ion's own stage times.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <vector>

#include "../../include/icarion_gpu.h"

namespace icarion {

// One waveform's value at t, from its descriptor and, for ARBITRARY, its
// n_points (time, value) pairs.  The host builds step tables with it and the
// RK45 kernels call it at each ion's stage times; the operation order is the
// reference's in every branch.
__host__ __device__ inline double waveform_eval(const icarion_waveform_desc& d, const double* ts, const double* vs,
                                                uint32_t n, double t) {
    constexpr double kTwoPi = 2.0 * 3.14159265358979323846;
    const double* p = d.p;
    switch (d.type) {
        case ICARION_WAVEFORM_CONSTANT:
            return p[0];
        case ICARION_WAVEFORM_LINEAR: {       // LinearWaveform::evaluate
            if (t < p[2]) return p[0];
            if (t >= p[3]) return d.clamp ? p[1] : p[0];
            const double duration = p[3] - p[2];
            const double t_norm = (t - p[2]) / duration;
            return p[0] + (p[1] - p[0]) * t_norm;
        }
        case ICARION_WAVEFORM_QUADRATIC:      // QuadraticWaveform::evaluate
            if (t < p[3] || t > p[4]) return p[0];
            return p[0] + p[1] * t + p[2] * t * t;
        case ICARION_WAVEFORM_EXPONENTIAL: {  // ExponentialWaveform::evaluate
            auto at = [p](double rel) { return p[0] + p[1] * exp(p[2] * rel); };
            if (t < p[3]) return at(0.0);
            if (t >= p[4]) return d.clamp ? at(p[4] - p[3]) : at(0.0);
            return at(t - p[3]);
        }
        case ICARION_WAVEFORM_SINUSOIDAL: {   // SinusoidalWaveform::evaluate
            const double omega = kTwoPi * p[2];
            return p[0] + p[1] * sin(omega * t + p[3]);
        }
        case ICARION_WAVEFORM_PWM: {          // PWMWaveform::evaluate
            if (p[2] <= 0.0) return p[0];
            auto at = [p](double t_rel) {
                const double period = 1.0 / p[2];
                const double phase_shift = p[4] / (kTwoPi * p[2]);
                double phase = fmod(t_rel + phase_shift, period);
                if (phase < 0.0) phase += period;
                return (phase < p[3] * period) ? p[1] : p[0];
            };
            if (t < p[5]) return p[0];
            if (t >= p[6]) return d.clamp ? at(p[6] - p[5]) : p[0];
            return at(t - p[5]);
        }
        case ICARION_WAVEFORM_PULSED:         // PulsedWaveform::evaluate
            return (t >= p[2] && t < p[2] + p[3]) ? p[1] : p[0];
        default:
            break;
    }
    // ArbitraryWaveform::evaluate (WaveformConfig.cpp:10-91)
    if (t <= ts[0]) return vs[0];
    if (t >= ts[n - 1]) return vs[n - 1];
    uint32_t lo = 0, cnt = n;   // std::lower_bound: the first ts[idx] >= t
    while (cnt > 0) {
        const uint32_t half = cnt / 2;
        if (ts[lo + half] < t) {
            lo += half + 1;
            cnt -= half + 1;
        } else {
            cnt = half;
        }
    }
    const uint32_t idx = lo;
    if (idx == 0) return vs[0];
    const uint32_t i0 = idx - 1, i1 = idx;
    const double t0 = ts[i0], t1 = ts[i1], v0 = vs[i0], v1 = vs[i1];
    if (d.interp == ICARION_INTERP_STEP) return v0;
    const double alpha = (t - t0) / (t1 - t0);
    if (d.interp != ICARION_INTERP_CUBIC) return v0 + alpha * (v1 - v0);
    double m0, m1;
    if (i0 == 0) m0 = (v1 - v0) / (t1 - t0);
    else m0 = (v1 - vs[i0 - 1]) / (t1 - ts[i0 - 1]);
    if (i1 == n - 1) m1 = (v1 - v0) / (t1 - t0);
    else m1 = (vs[i1 + 1] - v0) / (ts[i1 + 1] - t0);
    const double h00 = 2 * alpha * alpha * alpha - 3 * alpha * alpha + 1;
    const double h10 = alpha * alpha * alpha - 2 * alpha * alpha + alpha;
    const double h01 = -2 * alpha * alpha * alpha + 3 * alpha * alpha;
    const double h11 = alpha * alpha * alpha - alpha * alpha;
    const double dt = t1 - t0;
    return h00 * v0 + h10 * dt * m0 + h01 * v1 + h11 * dt * m1;
}

struct HostWaveform {
    icarion_waveform_desc d{};
    std::vector<double> ts, vs;

    double eval(double t) const {
        return waveform_eval(d, ts.data(), vs.data(), static_cast<uint32_t>(ts.size()), t);
    }
};

}  // namespace icarion
