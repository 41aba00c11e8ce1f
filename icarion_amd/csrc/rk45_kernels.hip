// rk45_kernels.hip - the RK45 engine step (RK45Strategy::step,
// RK45Strategy.cpp:354-520) on gfx950: FP64, and FP32 + Kahan stages with the
// error estimate, step control and time in FP64.
//
// Floating-point contraction is off in this file: the step-size controller
// accepts or rejects a trial on err <= 1, and separate multiply and add
// roundings (as in the reference and the oracle) keep those decisions the
// same as theirs instead of moving an error that sits within rounding of
// the threshold to the other side.
#pragma clang fp contract(off)

#include "waveform.hpp"
#include "step_device.hpp"
#include "step_device_f32.hpp"

namespace icarion {

// ======================= RK45 (one engine step per launch) ==================
//
// RK45Strategy::step (RK45Strategy.cpp:354-520) per ion, inside the engine step
// of SimulationEngine::process_timestep with per-ion dt: every live ion starts
// at the shared engine time T with its own trial dt, the accepted dt drives its
// collision and its time t_i, and the next engine time is max_i t_i
// (SimulationEngineUtils.cpp:49-83).  That max couples all ions once per step,
// so one launch is one engine step; rk45_advance_kernel forms the next engine
// time from the reductions this launch leaves in the Rk45Engine record.

namespace dp {
// Dormand-Prince 5(4) (RK45Strategy.cpp:101-157).  The propagated solution is
// the a7 row ("y4" in the reference), the embedded one the b5 weights ("y5").
constexpr double c2 = 1.0 / 5.0, c3 = 3.0 / 10.0, c4 = 4.0 / 5.0, c5 = 8.0 / 9.0;
constexpr double a21 = 1.0 / 5.0;
constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0,
                 a54 = -212.0 / 729.0;
constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                 a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0;
constexpr double a71 = 35.0 / 384.0, a73 = 500.0 / 1113.0, a74 = 125.0 / 192.0,
                 a75 = -2187.0 / 6784.0, a76 = 11.0 / 84.0;
constexpr double b51 = 5179.0 / 57600.0, b53 = 7571.0 / 16695.0, b54 = 393.0 / 640.0,
                 b55 = -92097.0 / 339200.0, b56 = 187.0 / 2100.0, b57 = 1.0 / 40.0;
constexpr double kBeta = 0.04, kAlpha = 0.2 - 0.04 * 0.75, kRejectExp = 0.2;
constexpr int kMaxReject = 10;
constexpr double kMinError = 1e-10, kDtMinTol = 1.001;
}  // namespace dp

// double -> unsigned bits with the same order (for atomic max / min)
__device__ __forceinline__ unsigned long long ord_bits(double x) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double ord_value(unsigned long long o) {
    const unsigned long long b = (o >> 63) ? (o & 0x7FFFFFFFFFFFFFFFull) : ~o;
    return __longlong_as_double(static_cast<long long>(b));
}

// U_eff(t) / r0^2 (AnalyticalFieldModel.cpp:226-241) at the ion's stage time.
__device__ __forceinline__ double rf_coeff(const StepParams& P, double t) {
    return P.has_rf ? (P.rf_U + P.rf_V * cos(P.rf_omega * t)) / P.r0_sq : 0.0;
}

// The engine step's gas density over the domain's (update_dynamic_environments
// at the engine time T, SimulationEngineLifecycle.cpp:34-47): 1 without a
// pressure waveform; a pressure <= 0 gives 0 (the reference throws there).
__device__ __forceinline__ double rk45_density_ratio(const StepParams& P, double T) {
    if (!P.pressure_wave) return 1.0;
    const icarion_waveform_desc& w = P.waves[ICARION_N_VOLTAGE_SLOTS];
    const double p = waveform_eval(w, w.times_s, w.values, w.n_points, T);
    return p > 0.0 ? p * P.pressure_scale : 0.0;
}

// A voltage slot at t: its waveform where one is set, else the constant.
__device__ __forceinline__ double wave_slot(const StepParams& P, int k, double t) {
    if (P.wave_mask & (1u << k)) {
        const icarion_waveform_desc& w = P.waves[k];
        return waveform_eval(w, w.times_s, w.values, w.n_points, t);
    }
    return P.wave_const[k];
}

// The analytic IMS / quadrupole field's inputs at t with waveforms
// (AnalyticalFieldModel.cpp:136-161, :219-248): E_z = axial(t) / L (the
// quadrupole's where |axial| > 1e-12), and the RF coefficient
// (quad(t) + V(t) cos(2 pi f(t) t)) / r0^2 where |V(t)| > 1e-12 - the values
// the fixed-step paths take from their step tables.
__device__ __forceinline__ void wave_field_inputs(const StepParams& P, double t, double& c, double& ez) {
    const double ax = wave_slot(P, ICARION_SLOT_DC_AXIAL, t);
    if (P.instrument == ICARION_INSTRUMENT_QUADRUPOLE) ez = fabs(ax) > 1e-12 ? ax / P.L : 0.0;
    else ez = ax / P.L;
    c = 0.0;
    if (P.has_rf) {
        const double V = wave_slot(P, ICARION_SLOT_RF_VOLTAGE, t);
        if (fabs(V) > 1e-12) {
            const double f = wave_slot(P, ICARION_SLOT_RF_FREQUENCY, t);
            const double U = wave_slot(P, ICARION_SLOT_DC_QUAD, t);
            const double om = 2.0 * kPi * f;
            c = (U + V * cos(om * t)) / P.r0_sq;
        }
    }
}

// RK45Strategy::compute_new_step (RK45Strategy.cpp:315-351)
__device__ __forceinline__ double rk45_new_step(const Rk45Cfg& c, double dt, double err, double dt_min,
                                                double dt_max, double prev_err) {
    double f = err > 1.0 ? c.safety * pow(1.0 / err, dp::kRejectExp)
                         : c.safety * pow(prev_err / err, dp::kBeta) * pow(1.0 / err, dp::kAlpha);
    f = fmax(f, c.max_dec);
    f = fmin(f, c.max_inc);
    return fmin(fmax(dt * f, dt_min), dt_max);
}

__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_min_f64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

template <int FIELD, int COLL>
__global__ void __launch_bounds__(256) rk45_step_kernel(const StepParams P, const IonPtrs S, const Rk45Ptrs R) {
    using namespace dp;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Rk45Engine* E = R.eng;
    const double T = E->T;
    const unsigned long long kstep = E->step;
    const double nrel = rk45_density_ratio(P, T);
    if (!(nrel > 0.0)) {   // no step: the host reports the engine time
        if (i == 0) E->bad_pressure = 1ull;
        return;
    }
    uint32_t pushed = 0, ncoll_launch = 0, rejected = 0, failed = 0;
    double tmax = -INFINITY, bmin = INFINITY, dtmax = 0.0;
    if (i < P.n_ions) {
        const double birth = S.birth ? S.birth[i] : -INFINITY;
        uint8_t flags = S.flags[i];
        bool active = flags & 1, born = (flags >> 1) & 1;
        double t_i = R.t[i], dt_i = R.dt[i];
        double death = S.death[i];
        double arrival = S.arrival[i];
        uint32_t nc = S.ncoll[i];
        // prepare_ions_for_integration (SimulationEngineLifecycle.cpp:86-107)
        if (!born && T >= birth) {
            born = true;
            active = true;
        }
        if (!born && birth > T) bmin = birth;
        if (active && born) {
            d3 pos{S.x[i], S.y[i], S.z[i]};
            d3 vel{S.vx[i], S.vy[i], S.vz[i]};
            if (!contains_f64(P, pos)) {
                active = false;
                death = T;
            } else {
                const double m = S.mass[i];
                const double q = S.charge[i];
                const uint32_t sidx = S.species[i];
                const SpeciesDev sp = P.species[sidx];
                const double inv_m = 1.0 / m;
                double lpre = 0.0;
                if constexpr (COLL == COLL_DAMPING) {
                    if (P.damping_model == 3) {
                        if (P.alpha_gas > 0.0 && P.n > 0.0 && P.v_th > 0.0 && P.m_gas > 0.0) {
                            const double mu = (m * P.m_gas) / (m + P.m_gas);
                            const double cs1 = kPi * q * sqrt(P.alpha_gas / (4.0 * kPi * kEpsilon0 * mu));
                            lpre = P.n * cs1 * P.v_th * P.m_gas / (P.m_gas + m);
                        }
                    } else {
                        const double K0 = sp.K0_cm2_Vs;
                        if (K0 > 0.0 && m > 0.0 && P.n > 0.0) {
                            const double K0t = K0 * 1e-4 * sqrt(kStpTemp / fmax(P.T, 1.0));
                            lpre = q / ((K0t * (kLoschmidt / P.n)) * m);
                        }
                    }
                    lpre *= nrel;
                }
                // analytic fields take the RF coefficient, field arrays the stage time;
                // the damping force relative to the gas velocity at the stage position
                auto acc = [&](d3 p, d3 v, double tt) {
                    double c = FIELD >= FIELD_ARRAY ? tt : rf_coeff(P, tt);
                    double ez = P.Ez_const;
                    if (FIELD == FIELD_LINEAR && P.wave_mask) wave_field_inputs(P, tt, c, ez);
                    return force_f64<FIELD, COLL, true>(P, p, v, c, ez, q, m, lpre) * inv_m;
                };
                const Rk45Cfg& c = P.rk45;
                const double dt_min = fmax(dt_i * c.min_factor, c.abs_min);
                const double dt_max = dt_i * c.max_factor;
                double last_err = R.last_err[i];
                bool fsal = R.fsal[i] != 0;
                d3 k1a_saved{R.k1x[i], R.k1y[i], R.k1z[i]};
                const d3 p0 = pos, v0 = vel;
                double h = dt_i, used = -1.0, dt_next = 0.0;
                for (int attempt = 0; attempt < kMaxReject; ++attempt) {
                    const d3 k1v = v0;
                    const d3 k1a = fsal ? k1a_saved : acc(p0, v0, T);
                    const d3 k2v = v0 + k1a * (h * a21);
                    const d3 k2a = acc(p0 + k1v * (h * a21), k2v, T + c2 * h);
                    const d3 k3v = v0 + (k1a * a31 + k2a * a32) * h;
                    const d3 k3a = acc(p0 + (k1v * a31 + k2v * a32) * h, k3v, T + c3 * h);
                    const d3 k4v = v0 + (k1a * a41 + k2a * a42 + k3a * a43) * h;
                    const d3 k4a = acc(p0 + (k1v * a41 + k2v * a42 + k3v * a43) * h, k4v, T + c4 * h);
                    const d3 k5v = v0 + (k1a * a51 + k2a * a52 + k3a * a53 + k4a * a54) * h;
                    const d3 k5a = acc(p0 + (k1v * a51 + k2v * a52 + k3v * a53 + k4v * a54) * h, k5v, T + c5 * h);
                    const d3 k6v = v0 + (k1a * a61 + k2a * a62 + k3a * a63 + k4a * a64 + k5a * a65) * h;
                    const d3 k6a = acc(p0 + (k1v * a61 + k2v * a62 + k3v * a63 + k4v * a64 + k5v * a65) * h,
                                       k6v, T + h);
                    const d3 p7 = p0 + (k1v * a71 + k3v * a73 + k4v * a74 + k5v * a75 + k6v * a76) * h;
                    const d3 v7 = v0 + (k1a * a71 + k3a * a73 + k4a * a74 + k5a * a75 + k6a * a76) * h;
                    const d3 k7a = acc(p7, v7, T + h);
                    const d3 p5 = p0 + (k1v * b51 + k3v * b53 + k4v * b54 + k5v * b55 + k6v * b56 + v7 * b57) * h;
                    const d3 v5 = v0 + (k1a * b51 + k3a * b53 + k4a * b54 + k5a * b55 + k6a * b56 + k7a * b57) * h;
                    // RK45Strategy::estimate_error (RK45Strategy.cpp:272-313)
                    const double y4[6] = {p7.x, p7.y, p7.z, v7.x, v7.y, v7.z};
                    const double y5[6] = {p5.x, p5.y, p5.z, v5.x, v5.y, v5.z};
                    const double y0[6] = {p0.x, p0.y, p0.z, v0.x, v0.y, v0.z};
                    double err = 0.0;
#pragma unroll
                    for (int c6 = 0; c6 < 6; ++c6)
                        err = fmax(err, fabs(y5[c6] - y4[c6]) / (c.atol + c.rtol * fabs(y0[c6])));
                    err = fmax(err, kMinError);
                    if (err <= 1.0 || (c.accept_at_dt_min && h <= dt_min * kDtMinTol)) {
                        pos = p7;
                        vel = v7;
                        k1a_saved = k7a;
                        fsal = true;
                        dt_next = rk45_new_step(c, h, err, dt_min, dt_max, last_err);
                        last_err = err;
                        used = h;
                        break;
                    }
                    h = rk45_new_step(c, h, err, dt_min, dt_max, last_err);
                    ++rejected;
                    fsal = false;
                }
                if (used < 0.0) {
                    failed = 1;
                } else {
                    pushed = 1;
                    dtmax = used;
                    const double z_before = p0.z;
                    if constexpr (COLL != COLL_NONE) {
                        PhiloxStream rng(P.seed, i, kstep);
                        // collision substeps of the ion's own dt (n_sub = 1: one stage)
                        const double dtc = P.n_sub == 1 ? used : used / static_cast<double>(P.n_sub);
                        const double dtn = COLL == COLL_DAMPING ? dtc : dtc * nrel;
                        const d3 ug = gas_velocity_f64(P, pos);
                        for (int s = 0; s < P.n_sub; ++s) {
                            if (collide_f64<COLL, true>(P, sp, sidx, m, lpre, vel, rng, dtn, ug)) {
                                ++nc;
                                ++ncoll_launch;
                            }
                        }
                    }
                    if (P.detect && isnan(arrival)) {
                        if (z_before < P.det_z && pos.z >= P.det_z)
                            arrival = T + used * (P.det_z - z_before) / (pos.z - z_before);
                    }
                    // finalize_ions_after_integration (SimulationEngineLifecycle.cpp:137-192)
                    if (!contains_f64(P, pos)) {
                        active = false;
                        death = T;
                    } else {
                        t_i = t_i + used;
                        dt_i = dt_next > 0.0 ? dt_next : (used > 0.0 ? used : E->dt_fallback);
                        if (!(isfinite(pos.x) && isfinite(pos.y) && isfinite(pos.z) && isfinite(vel.x) &&
                              isfinite(vel.y) && isfinite(vel.z))) {
                            active = false;
                            death = T;
                        }
                    }
                    S.x[i] = pos.x; S.y[i] = pos.y; S.z[i] = pos.z;
                    S.vx[i] = vel.x; S.vy[i] = vel.y; S.vz[i] = vel.z;
                }
                R.last_err[i] = last_err;
                R.fsal[i] = fsal ? 1 : 0;
                R.k1x[i] = k1a_saved.x; R.k1y[i] = k1a_saved.y; R.k1z[i] = k1a_saved.z;
            }
        }
        tmax = t_i;
        R.t[i] = t_i;
        R.dt[i] = dt_i;
        S.flags[i] = static_cast<uint8_t>((active ? 1 : 0) | (born ? 2 : 0));
        S.death[i] = death;
        S.arrival[i] = arrival;
        S.ncoll[i] = nc;
    }
    add_counters(P, pushed, ncoll_launch);
    tmax = wave_max_f64(tmax);
    bmin = wave_min_f64(bmin);
    dtmax = wave_max_f64(dtmax);
    const uint32_t rej = wave_sum_u32(rejected);
    const uint32_t fl = wave_sum_u32(failed);
    if ((threadIdx.x & 63) == 0) {
        if (tmax > -INFINITY) atomicMax(&E->tmax, ord_bits(tmax));
        if (bmin < INFINITY) atomicMin(&E->bmin, ord_bits(bmin));
        if (dtmax > 0.0) atomicMax(&E->dtmax, ord_bits(dtmax));
        if (rej) atomicAdd(&E->rejected, static_cast<unsigned long long>(rej));
        if (fl) atomicAdd(&E->failed, static_cast<unsigned long long>(fl));
    }
}

// The same engine step with FP32 stages: positions and velocities are FP32
// (hi, lo) pairs advanced by compensated sums, forces and stage combinations in
// FP32; the error estimate y5 - y4 = h sum_j (b5_j - a7_j) k_j is formed
// directly from the stages and accumulated in FP64 (so it does not suffer the
// cancellation of a difference of two FP32 solutions), and the controller,
// per-ion time and dt, stage times and RF phase stay FP64.
template <int FIELD, int COLL>
__global__ void __launch_bounds__(256) rk45_step_f32_kernel(const StepParams P, const IonPtrs S,
                                                            const Rk45Ptrs R) {
    using namespace dp;
    // error weights of the embedded pair (b5 - a7; b5_7 for the FSAL stage)
    constexpr double e1 = b51 - a71, e3 = b53 - a73, e4 = b54 - a74, e5 = b55 - a75, e6 = b56 - a76,
                     e7 = b57;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const FastConst& C = P.fc;
    Rk45Engine* E = R.eng;
    const double T = E->T;
    const unsigned long long kstep = E->step;
    const double nrel = rk45_density_ratio(P, T);
    if (!(nrel > 0.0)) {   // no step: the host reports the engine time
        if (i == 0) E->bad_pressure = 1ull;
        return;
    }
    uint32_t pushed = 0, ncoll_launch = 0, rejected = 0, failed = 0;
    double tmax = -INFINITY, bmin = INFINITY, dtmax = 0.0;
    if (i < P.n_ions) {
        const double birth = S.birth ? S.birth[i] : -INFINITY;
        uint8_t flags = S.flags[i];
        bool active = flags & 1, born = (flags >> 1) & 1;
        double t_i = R.t[i], dt_i = R.dt[i];
        double death = S.death[i];
        double arrival = S.arrival[i];
        uint32_t nc = S.ncoll[i];
        if (!born && T >= birth) {
            born = true;
            active = true;
        }
        if (!born && birth > T) bmin = birth;
        if (active && born) {
            const double X[6] = {S.x[i], S.y[i], S.z[i], S.vx[i], S.vy[i], S.vz[i]};
            float hi[6], lo[6];
#pragma unroll
            for (int c6 = 0; c6 < 6; ++c6) {
                hi[c6] = (float)X[c6];
                lo[c6] = (float)(X[c6] - (double)hi[c6]);
            }
            if (!contains_f32(P, C, hi[0], hi[1], hi[2])) {
                active = false;
                death = T;
            } else {
                const double md = S.mass[i];
                const double qd = S.charge[i];
                const uint32_t sidx = S.species[i];
                const SpeciesDev sp = P.species[sidx];
                const float m = (float)md;
                const float qm = (float)(qd / md);
                const float inv_m_plus = (float)(1.0 / (md + P.m_gas));
                float lpre = 0.0f;
                if constexpr (COLL == COLL_DAMPING) {
                    double l = 0.0;
                    if (P.damping_model == 3) {
                        if (P.alpha_gas > 0.0 && P.n > 0.0 && P.v_th > 0.0 && P.m_gas > 0.0) {
                            const double mu = (md * P.m_gas) / (md + P.m_gas);
                            const double cs1 = kPi * qd * sqrt(P.alpha_gas / (4.0 * kPi * kEpsilon0 * mu));
                            l = P.n * cs1 * P.v_th * P.m_gas / (P.m_gas + md);
                        }
                    } else {
                        const double K0 = sp.K0_cm2_Vs;
                        if (K0 > 0.0 && md > 0.0 && P.n > 0.0) {
                            const double K0t = K0 * 1e-4 * sqrt(kStpTemp / fmax(P.T, 1.0));
                            l = qd / ((K0t * (kLoschmidt / P.n)) * md);
                        }
                    }
                    lpre = (float)(l * nrel);
                }
                auto acc = [&](f3 p, f3 v, double tt) {
                    double c = FIELD >= FIELD_ARRAY ? tt : rf_coeff(P, tt);
                    double ez = P.Ez_const;
                    if (FIELD == FIELD_LINEAR && P.wave_mask) wave_field_inputs(P, tt, c, ez);
                    return accel_f32<FIELD, COLL, true>(P, C, p, v, c, ez, qm, lpre);
                };
                const Rk45Cfg& c = P.rk45;
                const double dt_min = fmax(dt_i * c.min_factor, c.abs_min);
                const double dt_max = dt_i * c.max_factor;
                double last_err = R.last_err[i];
                bool fsal = R.fsal[i] != 0;
                f3 k1a_saved{(float)R.k1x[i], (float)R.k1y[i], (float)R.k1z[i]};
                const f3 p0{hi[0], hi[1], hi[2]}, v0{hi[3], hi[4], hi[5]};
                const double y0[6] = {X[0], X[1], X[2], X[3], X[4], X[5]};
                double h = dt_i, used = -1.0, dt_next = 0.0;
                f3 dp7{0.f, 0.f, 0.f}, dv7{0.f, 0.f, 0.f}, k7a_keep{0.f, 0.f, 0.f};
                for (int attempt = 0; attempt < kMaxReject; ++attempt) {
                    const float hf = (float)h;
                    const f3 k1v = v0;
                    const f3 k1a = fsal ? k1a_saved : acc(p0, v0, T);
                    const f3 k2v = v0 + k1a * (hf * (float)a21);
                    const f3 k2a = acc(p0 + k1v * (hf * (float)a21), k2v, T + c2 * h);
                    const f3 k3v = v0 + (k1a * (float)a31 + k2a * (float)a32) * hf;
                    const f3 k3a = acc(p0 + (k1v * (float)a31 + k2v * (float)a32) * hf, k3v, T + c3 * h);
                    const f3 k4v = v0 + (k1a * (float)a41 + k2a * (float)a42 + k3a * (float)a43) * hf;
                    const f3 k4a = acc(p0 + (k1v * (float)a41 + k2v * (float)a42 + k3v * (float)a43) * hf, k4v,
                                       T + c4 * h);
                    const f3 k5v = v0 + (k1a * (float)a51 + k2a * (float)a52 + k3a * (float)a53 +
                                         k4a * (float)a54) * hf;
                    const f3 k5a = acc(p0 + (k1v * (float)a51 + k2v * (float)a52 + k3v * (float)a53 +
                                             k4v * (float)a54) * hf,
                                       k5v, T + c5 * h);
                    const f3 k6v = v0 + (k1a * (float)a61 + k2a * (float)a62 + k3a * (float)a63 +
                                         k4a * (float)a64 + k5a * (float)a65) * hf;
                    const f3 k6a = acc(p0 + (k1v * (float)a61 + k2v * (float)a62 + k3v * (float)a63 +
                                             k4v * (float)a64 + k5v * (float)a65) * hf,
                                       k6v, T + h);
                    const f3 dp = (k1v * (float)a71 + k3v * (float)a73 + k4v * (float)a74 + k5v * (float)a75 +
                                   k6v * (float)a76) * hf;
                    const f3 dv = (k1a * (float)a71 + k3a * (float)a73 + k4a * (float)a74 + k5a * (float)a75 +
                                   k6a * (float)a76) * hf;
                    const f3 v7 = v0 + dv;
                    const f3 k7a = acc(p0 + dp, v7, T + h);
                    // RK45Strategy::estimate_error (RK45Strategy.cpp:272-313) on
                    // y5 - y4 formed from the stages in FP64
                    const double ep[3] = {
                        h * (e1 * k1v.x + e3 * k3v.x + e4 * k4v.x + e5 * k5v.x + e6 * k6v.x + e7 * v7.x),
                        h * (e1 * k1v.y + e3 * k3v.y + e4 * k4v.y + e5 * k5v.y + e6 * k6v.y + e7 * v7.y),
                        h * (e1 * k1v.z + e3 * k3v.z + e4 * k4v.z + e5 * k5v.z + e6 * k6v.z + e7 * v7.z)};
                    const double ev[3] = {
                        h * (e1 * k1a.x + e3 * k3a.x + e4 * k4a.x + e5 * k5a.x + e6 * k6a.x + e7 * k7a.x),
                        h * (e1 * k1a.y + e3 * k3a.y + e4 * k4a.y + e5 * k5a.y + e6 * k6a.y + e7 * k7a.y),
                        h * (e1 * k1a.z + e3 * k3a.z + e4 * k4a.z + e5 * k5a.z + e6 * k6a.z + e7 * k7a.z)};
                    double err = 0.0;
#pragma unroll
                    for (int c3i = 0; c3i < 3; ++c3i) {
                        err = fmax(err, fabs(ep[c3i]) / (c.atol + c.rtol * fabs(y0[c3i])));
                        err = fmax(err, fabs(ev[c3i]) / (c.atol + c.rtol * fabs(y0[3 + c3i])));
                    }
                    err = fmax(err, kMinError);
                    if (err <= 1.0 || (c.accept_at_dt_min && h <= dt_min * kDtMinTol)) {
                        dp7 = dp;
                        dv7 = dv;
                        k7a_keep = k7a;
                        fsal = true;
                        dt_next = rk45_new_step(c, h, err, dt_min, dt_max, last_err);
                        last_err = err;
                        used = h;
                        break;
                    }
                    h = rk45_new_step(c, h, err, dt_min, dt_max, last_err);
                    ++rejected;
                    fsal = false;
                }
                if (used < 0.0) {
                    failed = 1;
                } else {
                    pushed = 1;
                    dtmax = used;
                    k1a_saved = k7a_keep;
                    const float z_before = hi[2], z_before_lo = lo[2];
                    kahan_add(hi[0], lo[0], dp7.x);
                    kahan_add(hi[1], lo[1], dp7.y);
                    kahan_add(hi[2], lo[2], dp7.z);
                    kahan_add(hi[3], lo[3], dv7.x);
                    kahan_add(hi[4], lo[4], dv7.y);
                    kahan_add(hi[5], lo[5], dv7.z);
                    if constexpr (COLL != COLL_NONE) {
                        PhiloxStream rng(P.seed, i, kstep);
                        f3 v{hi[3], hi[4], hi[5]};
                        const f3 vin = v;
                        // collision substeps of the ion's own dt (n_sub = 1: one stage)
                        const double dtc = P.n_sub == 1 ? used : used / static_cast<double>(P.n_sub);
                        const float kbt2_dt_over_m = (float)(2.0 * kBoltzmann * P.T / md * dtc);
                        uint32_t hits = 0;
                        const f3 ug = gas_velocity_f32(P, C, f3{hi[0], hi[1], hi[2]});
                        for (int s = 0; s < P.n_sub; ++s)
                            hits += collide_f32<COLL, true>(P, C, sp, sidx, m, inv_m_plus, lpre, kbt2_dt_over_m, v,
                                                            rng, (float)(dtc * nrel), (float)(P.n * nrel * dtc), ug)
                                        ? 1u : 0u;
                        if (COLL == COLL_HSS || COLL == COLL_HSS_MIX) {
                            if (hits) {
                                hi[3] = v.x; hi[4] = v.y; hi[5] = v.z;
                                lo[3] = 0.0f; lo[4] = 0.0f; lo[5] = 0.0f;
                                nc += hits;
                                ncoll_launch += hits;
                            }
                        } else {
                            kahan_add(hi[3], lo[3], v.x - vin.x);
                            kahan_add(hi[4], lo[4], v.y - vin.y);
                            kahan_add(hi[5], lo[5], v.z - vin.z);
                        }
                    }
                    if (P.detect && isnan(arrival)) {
                        const double zb = (double)z_before + (double)z_before_lo;
                        const double za = (double)hi[2] + (double)lo[2];
                        if (zb < P.det_z && za >= P.det_z) arrival = T + used * (P.det_z - zb) / (za - zb);
                    }
                    if (!contains_f32(P, C, hi[0], hi[1], hi[2])) {
                        active = false;
                        death = T;
                    } else {
                        t_i = t_i + used;
                        dt_i = dt_next > 0.0 ? dt_next : (used > 0.0 ? used : E->dt_fallback);
                        if (!(isfinite(hi[0]) && isfinite(hi[1]) && isfinite(hi[2]) && isfinite(hi[3]) &&
                              isfinite(hi[4]) && isfinite(hi[5]))) {
                            active = false;
                            death = T;
                        }
                    }
                    S.x[i] = (double)hi[0] + (double)lo[0];
                    S.y[i] = (double)hi[1] + (double)lo[1];
                    S.z[i] = (double)hi[2] + (double)lo[2];
                    S.vx[i] = (double)hi[3] + (double)lo[3];
                    S.vy[i] = (double)hi[4] + (double)lo[4];
                    S.vz[i] = (double)hi[5] + (double)lo[5];
                }
                R.last_err[i] = last_err;
                R.fsal[i] = fsal ? 1 : 0;
                R.k1x[i] = k1a_saved.x; R.k1y[i] = k1a_saved.y; R.k1z[i] = k1a_saved.z;
            }
        }
        tmax = t_i;
        R.t[i] = t_i;
        R.dt[i] = dt_i;
        S.flags[i] = static_cast<uint8_t>((active ? 1 : 0) | (born ? 2 : 0));
        S.death[i] = death;
        S.arrival[i] = arrival;
        S.ncoll[i] = nc;
    }
    add_counters(P, pushed, ncoll_launch);
    tmax = wave_max_f64(tmax);
    bmin = wave_min_f64(bmin);
    dtmax = wave_max_f64(dtmax);
    const uint32_t rej = wave_sum_u32(rejected);
    const uint32_t fl = wave_sum_u32(failed);
    if ((threadIdx.x & 63) == 0) {
        if (tmax > -INFINITY) atomicMax(&E->tmax, ord_bits(tmax));
        if (bmin < INFINITY) atomicMin(&E->bmin, ord_bits(bmin));
        if (dtmax > 0.0) atomicMax(&E->dtmax, ord_bits(dtmax));
        if (rej) atomicAdd(&E->rejected, static_cast<unsigned long long>(rej));
        if (fl) atomicAdd(&E->failed, static_cast<unsigned long long>(fl));
    }
}

// next_engine_time_after_step (SimulationEngineUtils.cpp:49-83), total_time_s
// unbounded; then clears the reductions for the next step.
__global__ void rk45_advance_kernel(Rk45Engine* E) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (E->bad_pressure) return;   // the step did not run: the engine time stays
    const double T = E->T;
    double Tn = T;
    if (E->tmax != 0ull) Tn = fmax(T, ord_value(E->tmax));
    if (!(Tn > T)) {
        if (E->bmin != ~0ull) Tn = ord_value(E->bmin);
        else if (E->dtmax != 0ull && ord_value(E->dtmax) > 0.0) Tn = T + ord_value(E->dtmax);
        else Tn = T + E->dt_fallback;
    }
    E->T = Tn;
    E->step = E->step + 1;
    E->tmax = 0ull;
    E->bmin = ~0ull;
    E->dtmax = 0ull;
}

// Ion times / dts left unset by the upload take the first step's t0 / dt.
__global__ void rk45_init_kernel(Rk45Ptrs R, uint32_t n, double t0, double dt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (isnan(R.t[i])) R.t[i] = t0;
    if (!(R.dt[i] > 0.0)) R.dt[i] = dt;
}

// ---- launch ----------------------------------------------------------------

using Rk45Fn = void (*)(const StepParams, const IonPtrs, const Rk45Ptrs);

template <int FIELD>
Rk45Fn pick_rk45_coll(int coll, int precision) {
    if (precision == 1) {
        switch (coll) {   // FP32 stages: no EHSS geometry (as the fixed-step FP32 path)
            case COLL_NONE: return rk45_step_f32_kernel<FIELD, COLL_NONE>;
            case COLL_HSS: return rk45_step_f32_kernel<FIELD, COLL_HSS>;
            case COLL_HSS_MIX: return rk45_step_f32_kernel<FIELD, COLL_HSS_MIX>;
            case COLL_DAMPING: return rk45_step_f32_kernel<FIELD, COLL_DAMPING>;
            default: return nullptr;
        }
    }
    switch (coll) {
        case COLL_NONE: return rk45_step_kernel<FIELD, COLL_NONE>;
        case COLL_HSS: return rk45_step_kernel<FIELD, COLL_HSS>;
        case COLL_HSS_MIX: return rk45_step_kernel<FIELD, COLL_HSS_MIX>;
        case COLL_EHSS: return rk45_step_kernel<FIELD, COLL_EHSS>;
        case COLL_EHSS_TAB: return rk45_step_kernel<FIELD, COLL_EHSS_TAB>;
        case COLL_EHSS_MIX: return rk45_step_kernel<FIELD, COLL_EHSS_MIX>;
        case COLL_DAMPING: return rk45_step_kernel<FIELD, COLL_DAMPING>;
        default: return nullptr;
    }
}

// One RK45 engine step (step kernel + engine-time advance) on the stream.
hipError_t launch_rk45_step(int field, int coll, int precision, const StepParams& P, const IonPtrs& S,
                            const Rk45Ptrs& R, uint32_t block, hipStream_t stream) {
    Rk45Fn fn = field == FIELD_ORBITRAP ? pick_rk45_coll<FIELD_ORBITRAP>(coll, precision)
                : field == FIELD_ARRAY  ? pick_rk45_coll<FIELD_ARRAY>(coll, precision)
                : field == FIELD_ARRAY_X ? pick_rk45_coll<FIELD_ARRAY_X>(coll, precision)
                                        : pick_rk45_coll<FIELD_LINEAR>(coll, precision);
    if (!fn) return hipErrorInvalidValue;
    const uint32_t grid = (P.n_ions + block - 1) / block;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(block), 0, stream, P, S, R);
    hipLaunchKernelGGL(rk45_advance_kernel, dim3(1), dim3(64), 0, stream, R.eng);
    return hipGetLastError();
}

hipError_t launch_rk45_init(const Rk45Ptrs& R, uint32_t n, double t0, double dt, hipStream_t stream) {
    hipLaunchKernelGGL(rk45_init_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, R, n, t0, dt);
    return hipGetLastError();
}

}  // namespace icarion
