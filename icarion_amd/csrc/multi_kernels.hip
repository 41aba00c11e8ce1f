// multi_kernels.hip - the fused ion step over several domains (FP64, and FP32
// + Kahan with the domain logic kept in FP64).
//
// A multi-domain instrument: every domain keeps its own geometry, analytic
// field and gas, read per lane from a parameter record in global memory (a
// wave may hold ions in several domains).  Per step and ion, the engine's
// domain logic (SimulationEngineLifecycle.cpp:72-195) around the same push and
// collision bodies as the single-domain kernels:
//   prepare:  birth; the domain = DomainManager::find_domain_index of the
//             position (DomainManager.cpp:66-109); none -> death at t;
//   push:     RK4 (RK4Strategy.cpp:181-361) or Boris (BorisStrategy.cpp:18-98)
//             in that domain's field, the domain fixed for the whole step;
//   collide:  with that domain's gas (SimulationEngineCollisionRuntime.cpp:
//             35-60 groups the ions by their step domain);
//   finalize: find_domain_index of the new position; none -> the next domain
//             when the ion left through the end face
//             (forward_axial_bridge_domain, DomainManager.cpp:163-174), else
//             death at t; found -> shared_boundary_handoff_domain
//             (DomainManager.cpp:176-198) may move the index on; NaN kill.
// An ion kept alive by the bridge is looked up again at the next step start.
#include "step_device.hpp"
#include "step_device_f32.hpp"

namespace icarion {

// DomainManager::find_domain_index: the spans (cylindrical domains sorted by
// z_min) from std::lower_bound(z_max < z) on while z >= z_min, with the
// radial prefilter on the global x, y; then the linear scan in domain order.
// With an Orbitrap domain only the linear scan.
__device__ __forceinline__ int find_domain(const MultiParams& M, d3 p) {
    if (!M.has_orbitrap && M.n_spans > 0) {
        int first = 0, count = M.n_spans;
        while (count > 0) {
            const int step = count / 2;
            const int it = first + step;
            if (M.spans[it].z_max < p.z) {
                first = it + 1;
                count -= step + 1;
            } else {
                count = step;
            }
        }
        for (int s = first; s < M.n_spans; ++s) {
            const DomainSpan sp = M.spans[s];
            if (p.z < sp.z_min) break;
            if (p.x * p.x + p.y * p.y > sp.radius * sp.radius) continue;
            if (sp.idx >= 0 && sp.idx < M.n_dom && contains_f64(M.dom[sp.idx], p)) return sp.idx;
        }
    }
    for (int d = 0; d < M.n_dom; ++d)
        if (contains_f64(M.dom[d], p)) return d;
    return -1;
}

// forward_axial_bridge_domain: the next domain when z reached this domain's end.
__device__ __forceinline__ int bridge_domain(const MultiParams& M, int cur, d3 p) {
    if (cur < 0 || cur + 1 >= M.n_dom) return -1;
    const StepParams& D = M.dom[cur];
    return (p.z >= D.origin[2] + D.L) ? cur + 1 : -1;
}

// shared_boundary_handoff_domain: still in the current domain but within the
// tolerance of its end face and inside its end aperture -> the next domain.
__device__ __forceinline__ int handoff_domain(const MultiParams& M, int cur, int found, d3 p) {
    if (cur < 0 || found != cur || cur + 1 >= M.n_dom) return -1;
    const StepParams& D = M.dom[cur];
    const double local_z = p.z - D.origin[2];
    if (local_z < D.L - M.handoff_tol) return -1;
    const double dx = p.x - D.origin[0];
    const double dy = p.y - D.origin[1];
    return (dx * dx + dy * dy <= M.ap_sq[cur]) ? cur + 1 : -1;
}

// The hoisted damping factor of the single-domain kernels for one domain's gas
// (DampingForce.cpp:150-177 / :269-271).
__device__ __forceinline__ double damping_pre_f64(const StepParams& D, const SpeciesDev& sp, double m, double q) {
    double lpre = 0.0;
    if (D.damping_model == 3) {
        if (D.alpha_gas > 0.0 && D.n > 0.0 && D.v_th > 0.0 && D.m_gas > 0.0) {
            const double mu = (m * D.m_gas) / (m + D.m_gas);
            const double cs1 = kPi * q * sqrt(D.alpha_gas / (4.0 * kPi * kEpsilon0 * mu));
            lpre = D.n * cs1 * D.v_th * D.m_gas / (D.m_gas + m);
        }
    } else {
        const double K0 = sp.K0_cm2_Vs;
        if (K0 > 0.0 && m > 0.0 && D.n > 0.0) {
            const double K0t = K0 * 1e-4 * sqrt(kStpTemp / fmax(D.T, 1.0));
            lpre = q / ((K0t * (kLoschmidt / D.n)) * m);
        }
    }
    return lpre;
}

// EXT: the instantiation that also covers domains with a uniform field array
// (GridFieldProvider, one unscaled term), collision substeps and the
// AxialParabolic gas flow (each domain's own, D.flow_parabolic)
template <int COLL, bool EXT>
__device__ __forceinline__ d3 domain_force(const StepParams& D, d3 p, d3 v, double c, double ez, double q, double m,
                                           double lpre) {
    if constexpr (EXT) {
        if (D.field_array) return force_f64<FIELD_ARRAY, COLL, EXT>(D, p, v, c, ez, q, m, lpre);
    }
    if (D.instrument == 3) return force_f64<FIELD_ORBITRAP, COLL, EXT>(D, p, v, c, ez, q, m, lpre);
    return force_f64<FIELD_LINEAR, COLL, EXT>(D, p, v, c, ez, q, m, lpre);
}

template <int INTEG, int COLL, bool EXT>
__device__ __forceinline__ void multi_ion(const MultiParams& M, const IonPtrs& S, const uint32_t i, uint32_t& pushed,
                                          uint32_t& ncoll_launch) {
    d3 pos{S.x[i], S.y[i], S.z[i]};
    d3 vel{S.vx[i], S.vy[i], S.vz[i]};
    const double m = S.mass[i];
    const double q = S.charge[i];
    const uint32_t sidx = S.species[i];
    // the species row of the step domain's table (cross sections and mobility
    // resolved for that domain's gas), loaded when the domain changes
    SpeciesDev sp = M.species[sidx];
    int sp_dom = -1;
    uint8_t flags = S.flags[i];
    bool active = flags & 1, born = (flags >> 1) & 1;
    const double birth = S.birth ? S.birth[i] : -INFINITY;
    double death = S.death[i];
    double arrival = S.arrival[i];
    uint32_t nc = S.ncoll[i];
    int cur = S.domain[i];
    const double inv_m = 1.0 / m;
    double lpre = 0.0;
    int lpre_dom = -1;
    int found = -1;
    bool checked = false;   // found = the domain of the current position
    const double dt = M.dt;
    const double half_dt = dt * 0.5;
    for (uint32_t k = 0; k < M.n_steps; ++k) {
        const double t = M.t_table[M.k0 + k].t;
        if (!born) {
            if (t >= birth) {
                born = true;
                active = true;
                checked = false;
            } else {
                continue;
            }
        }
        if (!active) break;
        const int dom = checked ? found : find_domain(M, pos);
        if (dom < 0) {
            active = false;
            death = t;
            break;
        }
        cur = dom;
        const StepParams& D = M.dom[dom];
        if (dom != sp_dom) {
            sp = D.species[sidx];
            sp_dom = dom;
        }
        if constexpr (COLL == COLL_DAMPING) {
            if (dom != lpre_dom) {
                lpre = damping_pre_f64(D, sp, m, q);
                lpre_dom = dom;
            }
        }
        const StepEntry te = D.table[M.k0 + k];
        const double z_before = pos.z;
        // the domain's gas density at the step start over its constant one
        const double lp = lpre * te.nrel;
        if constexpr (INTEG == 0) {
            const d3 p0 = pos, v0 = vel;
            const d3 k1v = v0;
            const d3 k1a = domain_force<COLL, EXT>(D, p0, v0, te.c0, te.e0, q, m, lp) * inv_m;
            const d3 k2v = v0 + k1a * half_dt;
            const d3 k2a = domain_force<COLL, EXT>(D, p0 + k1v * half_dt, k2v, te.ch, te.eh, q, m, lp) * inv_m;
            const d3 k3v = v0 + k2a * half_dt;
            const d3 k3a = domain_force<COLL, EXT>(D, p0 + k2v * half_dt, k3v, te.ch, te.eh, q, m, lp) * inv_m;
            const d3 k4v = v0 + k3a * dt;
            const d3 k4a = domain_force<COLL, EXT>(D, p0 + k3v * dt, k4v, te.c1, te.e1, q, m, lp) * inv_m;
            const double h6 = dt / 6.0;
            pos = p0 + (k1v + k2v * 2.0 + k3v * 2.0 + k4v) * h6;
            vel = v0 + (k1a + k2a * 2.0 + k3a * 2.0 + k4a) * h6;
        } else {
            const d3 a = domain_force<COLL, EXT>(D, pos, vel, te.c0, te.e0, q, m, lp) * inv_m;
            const d3 vm = vel + a * half_dt;
            d3 vp = vm;
            if (D.magnetic) {
                const double qm = q * inv_m;
                const d3 tv = d3{D.B[0], D.B[1], D.B[2]} * (qm * dt * 0.5);
                const double s_f = 2.0 / (1.0 + dot(tv, tv));
                const d3 sv = tv * s_f;
                const d3 vprime = vm + cross(vm, tv);
                vp = vm + cross(vprime, sv);
            }
            pos = pos + vp * dt;
            vel = vp + a * half_dt;
        }
        ++pushed;
        if constexpr (COLL != COLL_NONE) {
            PhiloxStream rng(M.seed, i, M.step0 + k);
            // collision substeps in the EXT instantiation (one stage otherwise)
            const int ns = EXT ? D.n_sub : 1;
            const double dtc = EXT ? dt / static_cast<double>(D.n_sub) : dt;
            // rates take the density ratio through dt, the OU kick through gamma
            const double dtn = COLL == COLL_DAMPING ? dtc : dtc * te.nrel;
            // the step domain's gas velocity at the pushed position
            const d3 ug = EXT ? gas_velocity_f64(D, pos) : d3{0.0, 0.0, 0.0};
            for (int s = 0; s < ns; ++s) {
                bool hit;
                if constexpr (COLL == COLL_HSS_MIX)   // per domain: its mixture, else its single gas
                    hit = D.n_mix > 0 ? collide_f64<COLL_HSS_MIX, EXT>(D, sp, sidx, m, lp, vel, rng, dtn, ug)
                                      : collide_f64<COLL_HSS, EXT>(D, sp, sidx, m, lp, vel, rng, dtn, ug);
                else if constexpr (COLL == COLL_EHSS_MIX)
                    hit = D.n_emix > 0 ? collide_f64<COLL_EHSS_MIX, EXT>(D, sp, sidx, m, lp, vel, rng, dtn, ug)
                                       : collide_f64<COLL_EHSS, EXT>(D, sp, sidx, m, lp, vel, rng, dtn, ug);
                else
                    hit = collide_f64<COLL, EXT>(D, sp, sidx, m, lp, vel, rng, dtn, ug);
                if (hit) {
                    ++nc;
                    ++ncoll_launch;
                }
            }
        }
        if (D.detect && isnan(arrival)) {
            if (z_before < D.det_z && pos.z >= D.det_z)
                arrival = t + dt * (D.det_z - z_before) / (pos.z - z_before);
        }
        found = find_domain(M, pos);
        if (found < 0) {
            const int b = bridge_domain(M, cur, pos);
            if (b < 0) {
                active = false;
                death = t;
                break;
            }
            cur = b;
            checked = false;
        } else {
            const int h = handoff_domain(M, cur, found, pos);
            cur = h >= 0 ? h : found;
            checked = true;
        }
        if (!(isfinite(pos.x) && isfinite(pos.y) && isfinite(pos.z) && isfinite(vel.x) &&
              isfinite(vel.y) && isfinite(vel.z))) {
            active = false;
            death = t;
            break;
        }
    }
    S.x[i] = pos.x; S.y[i] = pos.y; S.z[i] = pos.z;
    S.vx[i] = vel.x; S.vy[i] = vel.y; S.vz[i] = vel.z;
    S.flags[i] = static_cast<uint8_t>((active ? 1 : 0) | (born ? 2 : 0));
    S.death[i] = death;
    S.arrival[i] = arrival;
    S.ncoll[i] = nc;
    S.domain[i] = cur;
}

template <int COLL, bool EXT>
__device__ __forceinline__ f3 domain_accel_f32(const StepParams& D, f3 p, f3 v, double c, double ez, float qm,
                                               float lpre) {
    if constexpr (EXT) {
        if (D.field_array) return accel_f32<FIELD_ARRAY, COLL, EXT>(D, D.fc, p, v, c, ez, qm, lpre);
    }
    if (D.instrument == 3) return accel_f32<FIELD_ORBITRAP, COLL, EXT>(D, D.fc, p, v, c, ez, qm, lpre);
    return accel_f32<FIELD_LINEAR, COLL, EXT>(D, D.fc, p, v, c, ez, qm, lpre);
}

// The FP32 + Kahan body of the single-domain fast path over several domains:
// position and velocity as FP32 (hi, lo) pairs, forces and collisions in FP32
// with each domain's FP32 constants; the domain lookup, bridge and handoff
// in FP64 on hi + lo, as the FP64 kernel takes them.
template <int INTEG, int COLL, bool EXT>
__device__ __forceinline__ void multi_ion_f32(const MultiParams& M, const IonPtrs& S, const uint32_t i,
                                              uint32_t& pushed, uint32_t& ncoll_launch) {
    const double X[6] = {S.x[i], S.y[i], S.z[i], S.vx[i], S.vy[i], S.vz[i]};
    float hi[6], lo[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        hi[c] = (float)X[c];
        lo[c] = (float)(X[c] - (double)hi[c]);
    }
    auto pos64 = [&]() {
        return d3{(double)hi[0] + (double)lo[0], (double)hi[1] + (double)lo[1], (double)hi[2] + (double)lo[2]};
    };
    const double md = S.mass[i];
    const double qd = S.charge[i];
    const uint32_t sidx = S.species[i];
    SpeciesDev sp = M.species[sidx];   // the step domain's row, reloaded with the domain's constants
    const float m = (float)md;
    const float qm = (float)(qd / md);
    uint8_t flags = S.flags[i];
    bool active = flags & 1, born = (flags >> 1) & 1;
    const double birth = S.birth ? S.birth[i] : -INFINITY;
    double death = S.death[i];
    double arrival = S.arrival[i];
    uint32_t nc = S.ncoll[i];
    int cur = S.domain[i];
    // the ion's constants in its current domain's gas
    int const_dom = -1;
    float lpre = 0.0f, inv_m_plus = 0.0f, kbt2_dt_over_m = 0.0f;
    int found = -1;
    bool checked = false;
    const float dt = (float)M.dt;
    const float half_dt = 0.5f * dt;
    const float h6 = dt * (1.0f / 6.0f);
    for (uint32_t k = 0; k < M.n_steps; ++k) {
        const double t = M.t_table[M.k0 + k].t;
        if (!born) {
            if (t >= birth) {
                born = true;
                active = true;
                checked = false;
            } else {
                continue;
            }
        }
        if (!active) break;
        const int dom = checked ? found : find_domain(M, pos64());
        if (dom < 0) {
            active = false;
            death = t;
            break;
        }
        cur = dom;
        const StepParams& D = M.dom[dom];
        if (dom != const_dom) {
            sp = D.species[sidx];
            if constexpr (COLL == COLL_DAMPING) lpre = (float)damping_pre_f64(D, sp, md, qd);
            inv_m_plus = (float)(1.0 / (md + D.m_gas));
            kbt2_dt_over_m = (float)(2.0 * kBoltzmann * D.T / md * D.dt);
            const_dom = dom;
        }
        const StepEntry te = D.table[M.k0 + k];
        const float z_before = hi[2];
        const float z_before_lo = lo[2];
        const float nrel = (float)te.nrel;
        const float lp = lpre * nrel;
        const f3 p0{hi[0], hi[1], hi[2]};
        const f3 v0{hi[3], hi[4], hi[5]};
        f3 dpos, dvel;
        if constexpr (INTEG == 0) {
            const f3 k1a = domain_accel_f32<COLL, EXT>(D, p0, v0, te.c0, te.e0, qm, lp);
            const f3 k2v = v0 + k1a * half_dt;
            const f3 k2a = domain_accel_f32<COLL, EXT>(D, p0 + v0 * half_dt, k2v, te.ch, te.eh, qm, lp);
            const f3 k3v = v0 + k2a * half_dt;
            const f3 k3a = domain_accel_f32<COLL, EXT>(D, p0 + k2v * half_dt, k3v, te.ch, te.eh, qm, lp);
            const f3 k4v = v0 + k3a * dt;
            const f3 k4a = domain_accel_f32<COLL, EXT>(D, p0 + k3v * dt, k4v, te.c1, te.e1, qm, lp);
            dpos = (v0 + k2v * 2.0f + k3v * 2.0f + k4v) * h6;
            dvel = (k1a + k2a * 2.0f + k3a * 2.0f + k4a) * h6;
        } else {
            const f3 a = domain_accel_f32<COLL, EXT>(D, p0, v0, te.c0, te.e0, qm, lp);
            const f3 vm = v0 + a * half_dt;
            f3 rot{0.0f, 0.0f, 0.0f};
            if (D.magnetic) {
                const f3 tv = f3{(float)D.B[0], (float)D.B[1], (float)D.B[2]} * (qm * half_dt);
                const float s_f = 2.0f / (1.0f + dot(tv, tv));
                const f3 vprime = vm + cross(vm, tv);
                rot = cross(vprime, tv * s_f);
            }
            dpos = (vm + rot) * dt;
            dvel = a * dt + rot;
        }
        kahan_add(hi[0], lo[0], dpos.x);
        kahan_add(hi[1], lo[1], dpos.y);
        kahan_add(hi[2], lo[2], dpos.z);
        kahan_add(hi[3], lo[3], dvel.x);
        kahan_add(hi[4], lo[4], dvel.y);
        kahan_add(hi[5], lo[5], dvel.z);
        ++pushed;
        if constexpr (COLL != COLL_NONE) {
            PhiloxStream rng(M.seed, i, M.step0 + k);
            f3 v{hi[3], hi[4], hi[5]};
            const f3 vin = v;
            // collision substeps in the EXT instantiation (one stage otherwise)
            const int ns = EXT ? D.n_sub : 1;
            const float sub = EXT ? 1.0f / static_cast<float>(D.n_sub) : 1.0f;
            const float dtc = EXT ? static_cast<float>(M.dt / static_cast<double>(D.n_sub)) : dt;
            uint32_t hits = 0;
            const f3 ug = EXT ? gas_velocity_f32(D, D.fc, f3{hi[0], hi[1], hi[2]}) : f3{0.0f, 0.0f, 0.0f};
            for (int s = 0; s < ns; ++s) {
                bool hit;
                if constexpr (COLL == COLL_HSS_MIX)
                    hit = D.n_mix > 0 ? collide_f32<COLL_HSS_MIX, EXT>(D, D.fc, sp, sidx, m, inv_m_plus, lp,
                                                                       kbt2_dt_over_m * sub, v, rng, dtc * nrel,
                                                                       D.fc.n_dt * sub * nrel, ug)
                                      : collide_f32<COLL_HSS, EXT>(D, D.fc, sp, sidx, m, inv_m_plus, lp,
                                                                   kbt2_dt_over_m * sub, v, rng, dtc * nrel,
                                                                   D.fc.n_dt * sub * nrel, ug);
                else
                    hit = collide_f32<COLL, EXT>(D, D.fc, sp, sidx, m, inv_m_plus, lp, kbt2_dt_over_m * sub, v,
                                                 rng, dtc * nrel, D.fc.n_dt * sub * nrel, ug);
                hits += hit ? 1u : 0u;
            }
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
        if (D.detect && isnan(arrival)) {
            const double zb = (double)z_before + (double)z_before_lo;
            const double za = (double)hi[2] + (double)lo[2];
            if (zb < D.det_z && za >= D.det_z) arrival = t + M.dt * (D.det_z - zb) / (za - zb);
        }
        const d3 p_after = pos64();
        found = find_domain(M, p_after);
        if (found < 0) {
            const int b = bridge_domain(M, cur, p_after);
            if (b < 0) {
                active = false;
                death = t;
                break;
            }
            cur = b;
            checked = false;
        } else {
            const int hnd = handoff_domain(M, cur, found, p_after);
            cur = hnd >= 0 ? hnd : found;
            checked = true;
        }
        if (!(isfinite(hi[0]) && isfinite(hi[1]) && isfinite(hi[2]) && isfinite(hi[3]) && isfinite(hi[4]) &&
              isfinite(hi[5]))) {
            active = false;
            death = t;
            break;
        }
    }
    S.x[i] = (double)hi[0] + (double)lo[0];
    S.y[i] = (double)hi[1] + (double)lo[1];
    S.z[i] = (double)hi[2] + (double)lo[2];
    S.vx[i] = (double)hi[3] + (double)lo[3];
    S.vy[i] = (double)hi[4] + (double)lo[4];
    S.vz[i] = (double)hi[5] + (double)lo[5];
    S.flags[i] = static_cast<uint8_t>((active ? 1 : 0) | (born ? 2 : 0));
    S.death[i] = death;
    S.arrival[i] = arrival;
    S.ncoll[i] = nc;
    S.domain[i] = cur;
}

template <int INTEG, int COLL, bool EXT>
__global__ void __launch_bounds__(256) step_multi_f32_kernel(const MultiParams M, const IonPtrs S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t pushed = 0, ncoll_launch = 0;
    if (i < M.n_ions) multi_ion_f32<INTEG, COLL, EXT>(M, S, i, pushed, ncoll_launch);
    add_counters_to(M.counters, M.counter_stripes, pushed, ncoll_launch);
}

template <int INTEG, int COLL, bool EXT>
__global__ void __launch_bounds__(256) step_multi_f64_kernel(const MultiParams M, const IonPtrs S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t pushed = 0, ncoll_launch = 0;
    if (i < M.n_ions) multi_ion<INTEG, COLL, EXT>(M, S, i, pushed, ncoll_launch);
    add_counters_to(M.counters, M.counter_stripes, pushed, ncoll_launch);
}

// One multi-domain kernel; ext: collision substeps or a domain with a field array
template <int INTEG, bool F32, int COLL>
static void launch_multi_kernel(bool ext, dim3 grid, uint32_t block, hipStream_t s, const MultiParams& M,
                                const IonPtrs& S) {
    if constexpr (F32) {
        if (ext)
            hipLaunchKernelGGL((step_multi_f32_kernel<INTEG, COLL, true>), grid, dim3(block), 0, s, M, S);
        else
            hipLaunchKernelGGL((step_multi_f32_kernel<INTEG, COLL, false>), grid, dim3(block), 0, s, M, S);
    } else {
        if (ext)
            hipLaunchKernelGGL((step_multi_f64_kernel<INTEG, COLL, true>), grid, dim3(block), 0, s, M, S);
        else
            hipLaunchKernelGGL((step_multi_f64_kernel<INTEG, COLL, false>), grid, dim3(block), 0, s, M, S);
    }
}

template <int INTEG, bool F32>
static hipError_t launch_multi_integ(int coll, bool ext, const MultiParams& M, const IonPtrs& S, uint32_t block,
                                     hipStream_t s) {
    const dim3 grid((M.n_ions + block - 1) / block);
    switch (coll) {
        case COLL_NONE: launch_multi_kernel<INTEG, F32, COLL_NONE>(ext, grid, block, s, M, S); break;
        case COLL_HSS: launch_multi_kernel<INTEG, F32, COLL_HSS>(ext, grid, block, s, M, S); break;
        case COLL_DAMPING: launch_multi_kernel<INTEG, F32, COLL_DAMPING>(ext, grid, block, s, M, S); break;
        case COLL_HSS_MIX:   // HSS with per-domain gas mixtures
            launch_multi_kernel<INTEG, F32, COLL_HSS_MIX>(ext, grid, block, s, M, S);
            break;
        case COLL_EHSS_MIX:   // geometry EHSS with per-domain gas mixtures, FP64 only
            if constexpr (F32) return hipErrorInvalidValue;
            else launch_multi_kernel<INTEG, false, COLL_EHSS_MIX>(ext, grid, block, s, M, S);
            break;
        case COLL_EHSS:   // geometry EHSS with each domain's gas, FP64 only
            if constexpr (F32) return hipErrorInvalidValue;
            else launch_multi_kernel<INTEG, false, COLL_EHSS>(ext, grid, block, s, M, S);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// integrator 0 RK4, 1 Boris; precision 0 FP64, 1 FP32 + Kahan;
// hipErrorInvalidValue for a collision kind with no multi-domain kernel.
hipError_t launch_multi_step(int integrator, int coll, int precision, const MultiParams& M, const IonPtrs& S,
                             uint32_t block, hipStream_t s) {
    if (M.n_ions == 0) return hipSuccess;
    const bool f32 = precision == 1;
    const bool ext = M.n_sub > 1 || M.has_field_array || M.has_flow;
    if (integrator == 0) return f32 ? launch_multi_integ<0, true>(coll, ext, M, S, block, s)
                                    : launch_multi_integ<0, false>(coll, ext, M, S, block, s);
    if (integrator == 1) return f32 ? launch_multi_integ<1, true>(coll, ext, M, S, block, s)
                                    : launch_multi_integ<1, false>(coll, ext, M, S, block, s);
    return hipErrorInvalidValue;
}

}  // namespace icarion
