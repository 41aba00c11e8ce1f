// waveform.hpp - host evaluation of time-varying voltages (ValueOrWaveform:
// the waveform types of WaveformConfig.h:19-200 and the interpolated table of
// WaveformConfig.cpp:10-91).  The host evaluates them at every stage time into
// the FP64 step table; kernels never see a waveform.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/icarion_gpu.h"

namespace icarion {

struct HostWaveform {
    icarion_waveform_desc d{};
    std::vector<double> ts, vs;

    double eval(double t) const {
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
                auto at = [p](double rel) { return p[0] + p[1] * std::exp(p[2] * rel); };
                if (t < p[3]) return at(0.0);
                if (t >= p[4]) return d.clamp ? at(p[4] - p[3]) : at(0.0);
                return at(t - p[3]);
            }
            case ICARION_WAVEFORM_SINUSOIDAL: {   // SinusoidalWaveform::evaluate
                const double omega = kTwoPi * p[2];
                return p[0] + p[1] * std::sin(omega * t + p[3]);
            }
            case ICARION_WAVEFORM_PWM: {          // PWMWaveform::evaluate
                if (p[2] <= 0.0) return p[0];
                auto at = [p](double t_rel) {
                    const double period = 1.0 / p[2];
                    const double phase_shift = p[4] / (kTwoPi * p[2]);
                    double phase = std::fmod(t_rel + phase_shift, period);
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
                return arbitrary(t);
        }
    }

    // ArbitraryWaveform::evaluate (WaveformConfig.cpp:10-91)
    double arbitrary(double t) const {
        if (t <= ts.front()) return vs.front();
        if (t >= ts.back()) return vs.back();
        const size_t idx = static_cast<size_t>(std::lower_bound(ts.begin(), ts.end(), t) - ts.begin());
        if (idx == 0) return vs[0];
        const size_t i0 = idx - 1, i1 = idx;
        const double t0 = ts[i0], t1 = ts[i1], v0 = vs[i0], v1 = vs[i1];
        if (d.interp == ICARION_INTERP_STEP) return v0;
        const double alpha = (t - t0) / (t1 - t0);
        if (d.interp != ICARION_INTERP_CUBIC) return v0 + alpha * (v1 - v0);
        double m0, m1;
        if (i0 == 0) m0 = (v1 - v0) / (t1 - t0);
        else m0 = (v1 - vs[i0 - 1]) / (t1 - ts[i0 - 1]);
        if (i1 == ts.size() - 1) m1 = (v1 - v0) / (t1 - t0);
        else m1 = (vs[i1 + 1] - v0) / (ts[i1 + 1] - t0);
        const double h00 = 2 * alpha * alpha * alpha - 3 * alpha * alpha + 1;
        const double h10 = alpha * alpha * alpha - 2 * alpha * alpha + alpha;
        const double h01 = -2 * alpha * alpha * alpha + 3 * alpha * alpha;
        const double h11 = alpha * alpha * alpha - alpha * alpha;
        const double dt = t1 - t0;
        return h00 * v0 + h10 * dt * m0 + h01 * v1 + h11 * dt * m1;
    }
};

}  // namespace icarion
