// step_kernels.hip - fused multi-step ion kernels for gfx950.
//
// One ion per lane.  A lane loads its ion once (coalesced SoA), runs
// params.n_steps engine steps with the state in registers, and stores it
// once; per step it does what the reference engine does to one ion:
//   prepare  (SimulationEngineLifecycle.cpp:86-107): birth, domain lookup;
//   push     (RK4Strategy.cpp:181-361 or BorisStrategy.cpp:18-98);
//   collide  (HSS: HSSCollisionHandler.cpp:257-376, EHSS:
//             EHSSCollisionHandler.cpp:222-455, OU kick: CollisionKernels.cpp:171-214);
//   finalize (SimulationEngineLifecycle.cpp:137-192): boundary elimination with
//             death at the step-start time, NaN kill.
// Ions never interact, so running these phases back to back per ion gives
// the reference's per-ion result.  Per-step wave-uniform values (engine time,
// RF coefficient) come from a table read with scalar loads.  Counters are
// reduced per wavefront before one atomic per wave.
#include "step_device.hpp"
#include "step_device_f32.hpp"
#include "surrogate.hpp"

namespace icarion {

// SUB: collision substeps; FLOW: the AxialParabolic gas flow (both read from P
// at run time, instantiated apart because they cost the plain kernels
// registers).
template <int INTEG, int FIELD, int COLL, bool SUB = false, bool FLOW = false>
__device__ __forceinline__ void f64_ion(const StepParams& P, const IonPtrs& S, const uint32_t i,
        uint32_t& pushed, uint32_t& ncoll_launch) {
    d3 pos{S.x[i], S.y[i], S.z[i]};
    d3 vel{S.vx[i], S.vy[i], S.vz[i]};
    const double m = S.mass[i];
    const double q = S.charge[i];
    const uint32_t sidx = S.species[i];
    const SpeciesDev sp = P.species[sidx];
    uint8_t flags = S.flags[i];
    bool active = flags & 1, born = (flags >> 1) & 1;
    const double birth = S.birth ? S.birth[i] : -INFINITY;
    double death = S.death[i];
    double arrival = S.arrival[i];
    uint32_t nc = S.ncoll[i];
    const double inv_m = 1.0 / m;
    // hoisted damping factor (DampingForce.cpp:150-177 / :269-271)
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
    }
    bool checked = false;   // position known to be inside the domain
    const double dt = P.dt;
    const double half_dt = dt * 0.5;
    for (uint32_t k = 0; k < P.n_steps; ++k) {
        const StepEntry te = P.table[k];
        const double t = te.t;
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
        if (!checked && !contains_f64(P, pos)) {
            active = false;
            death = t;
            break;
        }
        const double z_before = pos.z;
        // the step's gas density (update_dynamic_environments at the step
        // start, SimulationEngine.cpp:380): the damping coefficient and the
        // collision rates scale with it
        const double lp = lpre * te.nrel;
        if constexpr (INTEG == 0) {
            // RK4Strategy::step_batch_with_stage_refresh (RK4Strategy.cpp:181-361)
            const d3 p0 = pos, v0 = vel;
            const d3 k1v = v0;
            const d3 k1a = force_f64<FIELD, COLL, FLOW>(P, p0, v0, te.c0, te.e0, q, m, lp) * inv_m;
            const d3 k2v = v0 + k1a * half_dt;
            const d3 k2a = force_f64<FIELD, COLL, FLOW>(P, p0 + k1v * half_dt, k2v, te.ch, te.eh, q, m, lp) * inv_m;
            const d3 k3v = v0 + k2a * half_dt;
            const d3 k3a = force_f64<FIELD, COLL, FLOW>(P, p0 + k2v * half_dt, k3v, te.ch, te.eh, q, m, lp) * inv_m;
            const d3 k4v = v0 + k3a * dt;
            const d3 k4a = force_f64<FIELD, COLL, FLOW>(P, p0 + k3v * dt, k4v, te.c1, te.e1, q, m, lp) * inv_m;
            const double h6 = dt / 6.0;
            pos = p0 + (k1v + k2v * 2.0 + k3v * 2.0 + k4v) * h6;
            vel = v0 + (k1a + k2a * 2.0 + k3a * 2.0 + k4a) * h6;
        } else {
            // BorisStrategy::step (BorisStrategy.cpp:18-98)
            const d3 a = force_f64<FIELD, COLL, FLOW>(P, pos, vel, te.c0, te.e0, q, m, lp) * inv_m;
            const d3 vm = vel + a * half_dt;
            d3 vp = vm;
            if (P.magnetic) {
                const double qm = q * inv_m;
                const d3 tv = d3{P.B[0], P.B[1], P.B[2]} * (qm * dt * 0.5);
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
            PhiloxStream rng(P.seed, i, P.step0 + k);
            // collision substeps, each an event of its own, in the SUB
            // instantiation (the loop costs the one-stage kernels registers)
            const int ns = SUB ? P.n_sub : 1;
            const double dtc = SUB ? dt / static_cast<double>(P.n_sub) : dt;
            // a rate n sigma |v| dt takes the density ratio through dt; the OU
            // kick's variance 2 gamma kT dt / m through gamma alone
            const double dtn = COLL == COLL_DAMPING ? dtc : dtc * te.nrel;
            // the gas velocity at the pushed position, for every substep
            // (SimulationEngineCollisionRuntime.cpp:122-130)
            const d3 ug = FLOW ? gas_velocity_f64(P, pos) : d3{0.0, 0.0, 0.0};
            for (int s = 0; s < ns; ++s) {
                if (collide_f64<COLL, FLOW>(P, sp, sidx, m, lp, vel, rng, dtn, ug)) {
                    ++nc;
                    ++ncoll_launch;
                }
            }
        }
        if (P.detect && isnan(arrival)) {
            if (z_before < P.det_z && pos.z >= P.det_z)
                arrival = t + dt * (P.det_z - z_before) / (pos.z - z_before);
        }
        if (!contains_f64(P, pos)) {
            active = false;
            death = t;
            break;
        }
        checked = true;
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
}

template <int INTEG, int FIELD, int COLL, bool SUB = false>
__global__ void __launch_bounds__(256) step_f64_kernel(const StepParams P, const IonPtrs S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t pushed = 0, ncoll_launch = 0;
    if (i < P.n_ions) f64_ion<INTEG, FIELD, COLL, SUB, SUB>(P, S, P.order ? P.order[i] : i, pushed, ncoll_launch);
    add_counters(P, pushed, ncoll_launch);
}

// ======================= FP64 geometry EHSS, wave-cooperative ===============
//
// The plain EHSS method of ehss_f64 / ehss_core_f64 with the work of one
// collision spread over the whole wavefront.  Each lane pushes its own ion and
// runs its own collision test (the handler's Maxwellian neutral and
// P = 1 - exp(-n sigma_eff |v_rel| dt)); then the wavefront takes the ions that
// collide one after another, and for each of them:
//   * its random stream is handed to every lane (the stream is counter-based,
//     so every lane draws the same numbers the ion's own lane would);
//   * atom j of its cluster is handled by lane j mod 64: b_max is a wave max
//     over the rotated atoms' radial extents, and for each impact parameter
//     the first contact is a wave min over (s_hit, j) - the smallest s_hit,
//     the lowest atom index among equal ones, which is the sequential loop's
//     choice (CollisionKernels.cpp:283-381);
//   * every lane forms the contact normal of the winning atom with the
//     sequential loop's arithmetic, and the reflected velocity; the ion's lane
//     keeps it.
// So the wavefront spends the atoms x attempts work of the collisions it has,
// instead of waiting on its slowest lane while the lanes without a collision
// idle.

__device__ __forceinline__ uint32_t readlane_u32(uint32_t x, int lane) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), lane));
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ d3 readlane_d3(d3 v, int lane) {
    return d3{readlane_f64(v.x, lane), readlane_f64(v.y, lane), readlane_f64(v.z, lane)};
}

__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// (s, j) <- the lexicographic minimum over the wavefront, on every lane
__device__ __forceinline__ void wave_min_contact(double& s, uint32_t& j) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double so = __shfl_xor(s, off, 64);
        const uint32_t jo = __shfl_xor(j, off, 64);
        if (so < s || (so == s && jo < j)) {
            s = so;
            j = jo;
        }
    }
}

// Rotated atoms of the collision in progress, per wavefront: the first
// kWaveAtoms atoms of a cluster (x, y, z in the collision frame and
// (r_atom + r_gas)^2), each written and read by its own lane; atoms past that
// are rotated again from the table on every impact parameter.
constexpr uint32_t kWaveAtoms = 512;

// A per-(ion, step) stream whose state is wave-uniform: the draws of one ion
// replayed by every lane, with its state and the Philox rounds held in scalar
// registers (the same sequence as PhiloxStream).
struct WaveStream {
    uint32_t k0, k1, ion, s_lo, s_hi, block;
    uint32_t buf[4];
    int pos;

    __device__ __forceinline__ void refill() {
        const u32x4 o = philox4x32_10(__builtin_amdgcn_readfirstlane(block), __builtin_amdgcn_readfirstlane(ion),
                                      __builtin_amdgcn_readfirstlane(s_lo), __builtin_amdgcn_readfirstlane(s_hi),
                                      __builtin_amdgcn_readfirstlane(k0), __builtin_amdgcn_readfirstlane(k1));
#pragma unroll
        for (int c = 0; c < 4; ++c) buf[c] = __builtin_amdgcn_readfirstlane(o.v[c]);
        ++block;
        pos = 0;
    }
    // pos is 0 or 2: the words are selected, not indexed, so that the block
    // stays in (scalar) registers instead of a private-memory array
    __device__ __forceinline__ double uniform_f64() {
        if (pos >= 4) refill();
        const bool second = pos != 0;
        const double u = uniform53(second ? buf[2] : buf[0], second ? buf[3] : buf[1]);
        pos += 2;
        return u;
    }
};

// ehss_f64 of one ion, called by every lane of the wavefront with the same
// (wave-uniform) arguments; returns the ion's new velocity on every lane.
// lds: this wavefront's kWaveAtoms-entry staging area.
__device__ d3 ehss_wave_f64(const StepParams& P, d3 v_ion, d3 v_n, double mi, double mn, double r_gas,
                            uint32_t a0, uint32_t nat, WaveStream& rng, uint32_t lane, double4* lds) {
    const double u1 = rng.uniform_f64();
    const double u2 = rng.uniform_f64();
    const double u3 = rng.uniform_f64();
    if (nat == 0) return v_ion;
    double s2, c2, s3, c3;
    sincos(2.0 * kPi * u2, &s2, &c2);
    sincos(2.0 * kPi * u3, &s3, &c3);
    const double q1 = sqrt(1.0 - u1) * s2;
    const double q2 = sqrt(1.0 - u1) * c2;
    const double q3 = sqrt(u1) * s3;
    const double q4 = sqrt(u1) * c3;
    const double x = q1, y = q2, z = q3, w = q4;
    double O[9];
    O[0] = 1 - 2 * (y * y + z * z); O[1] = 2 * (x * y - z * w);     O[2] = 2 * (x * z + y * w);
    O[3] = 2 * (x * y + z * w);     O[4] = 1 - 2 * (x * x + z * z); O[5] = 2 * (y * z - x * w);
    O[6] = 2 * (x * z - y * w);     O[7] = 2 * (y * z + x * w);     O[8] = 1 - 2 * (x * x + y * y);
    d3 vrel = v_ion - v_n;
    double vrel_mag = sqrt(dot(vrel, vrel));
    if (vrel_mag <= kMinVelocityMag) {
        // iso_dir_f64 on the wave-uniform stream
        const double w1 = rng.uniform_f64();
        const double w2 = rng.uniform_f64();
        const double cosT = 2.0 * w1 - 1.0;
        const double sinT = sqrt(fmax(0.0, 1.0 - cosT * cosT));
        double sp, cp;
        sincos(2.0 * kPi * w2, &sp, &cp);
        vrel = d3{sinT * cp, sinT * sp, cosT} * kMinVelocityMag;
        vrel_mag = kMinVelocityMag;
    }
    const d3 ax{vrel.x / vrel_mag, vrel.y / vrel_mag, vrel.z / vrel_mag};
    d3 t1, t2;
    ortho_basis_f64(ax, t1, t2);
    const double Bm[3][3] = {{t1.x, t2.x, ax.x}, {t1.y, t2.y, ax.y}, {t1.z, t2.z, ax.z}};
    double Rm[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            Rm[3 * a + b] = Bm[a][0] * O[b] + Bm[a][1] * O[3 + b] + Bm[a][2] * O[6 + b];
    const float4* __restrict__ atoms = P.atoms + a0;
    double bm = 0.0;
    for (uint32_t j = lane; j < nat; j += 64) {
        const float4 f = atoms[j];
        const d3 ra = mat_mul(Rm, d3{f.x, f.y, f.z});
        const double sra = dot(ax, ra);
        const d3 rp = ra - ax * sra;
        const double cur = sqrt(dot(rp, rp)) + static_cast<double>(f.w);
        bm = fmax(bm, cur);
        if (j < kWaveAtoms) {
            const double Rsum = static_cast<double>(f.w) + r_gas;
            lds[j] = make_double4(ra.x, ra.y, ra.z, Rsum * Rsum);
        }
    }
    double b_max = wave_max_f64(bm) + r_gas;
    for (int attempt = 0; attempt < kEhssMaxAttempts; ++attempt) {
        const double uA = rng.uniform_f64();
        const double uB = rng.uniform_f64();
        const double b = b_max * sqrt(uA);
        double sphi, cphi;
        sincos(2.0 * kPi * uB, &sphi, &cphi);
        const d3 off = t1 * (b * cphi) + t2 * (b * sphi);
        double best_s = INFINITY;
        uint32_t best_j = 0xFFFFFFFFu;
        for (uint32_t j = lane; j < nat; j += 64) {
            d3 ra;
            double Rsum2;
            if (j < kWaveAtoms) {
                const double4 a = lds[j];
                ra = d3{a.x, a.y, a.z};
                Rsum2 = a.w;
            } else {
                const float4 f = atoms[j];
                ra = mat_mul(Rm, d3{f.x, f.y, f.z});
                const double Rsum = static_cast<double>(f.w) + r_gas;
                Rsum2 = Rsum * Rsum;
            }
            const d3 rel = off - ra;
            const double sstar = dot(ax, rel);
            const d3 dv = rel - ax * sstar;
            const double dmin2 = dot(dv, dv);
            if (dmin2 <= Rsum2) {
                const double s_hit = sstar - sqrt(fmax(0.0, Rsum2 - dmin2));
                if (s_hit < best_s) {
                    best_s = s_hit;
                    best_j = j;
                }
            }
        }
        // no lane found a contact: the next impact parameter (wave-uniform)
        if (__ballot(best_j != 0xFFFFFFFFu) == 0ull) {
            if (attempt == kBmaxExpansionAttempt) b_max *= kBmaxExpansionFactor;
            continue;
        }
        wave_min_contact(best_s, best_j);
        {
            // the contact normal of the first atom hit, as the sequential loop forms it
            const float4 f = atoms[best_j];
            const d3 ra = mat_mul(Rm, d3{f.x, f.y, f.z});
            const double Rsum = static_cast<double>(f.w) + r_gas;
            const double Rsum2 = Rsum * Rsum;
            const d3 rel = off - ra;
            const double sstar = dot(ax, rel);
            const d3 dv = rel - ax * sstar;
            const double dmin2 = dot(dv, dv);
            const double h = sqrt(fmax(0.0, Rsum2 - dmin2));
            const double s_hit = sstar - h;
            const d3 ph = rel - ax * s_hit;
            const double phr2 = dot(ph, ph);
            d3 nb;
            if (phr2 <= kMinContactDistSq) {
                nb = ax * -1.0;
            } else {
                nb = ph * (1.0 / sqrt(phr2));
                if (dot(ax, nb) > 0.0) nb = nb * -1.0;
            }
            const double vdn = dot(vrel, nb);
            const d3 vr = vrel - nb * (2.0 * vdn);
            const double inv_mt = 1.0 / (mi + mn);
            const d3 Vcom = (v_ion * mi + v_n * mn) * inv_mt;
            return Vcom + vr * (mn * inv_mt);
        }
    }
    return v_ion;
}

// step_f64_kernel<INTEG, FIELD, COLL_EHSS> with the geometry method run by
// the wavefront (above).  Every lane stays in the step loop until no lane of
// its wavefront has an ion left to step (lanes past the last ion, dead ions
// and ions done are carried along), so the wave-wide parts see all 64 lanes.
template <int INTEG, int FIELD>
__global__ void __launch_bounds__(256) step_f64_ehss_kernel(const StepParams P, const IonPtrs S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const bool have = i < P.n_ions;
    d3 pos{0.0, 0.0, 0.0}, vel{0.0, 0.0, 0.0};
    double m = 1.0, q = 0.0, birth = -INFINITY, death = -1.0, arrival = NAN;
    uint32_t sidx = 0, nc = 0;
    bool active = false, born = true;
    if (have) {
        pos = d3{S.x[i], S.y[i], S.z[i]};
        vel = d3{S.vx[i], S.vy[i], S.vz[i]};
        m = S.mass[i];
        q = S.charge[i];
        sidx = S.species[i];
        const uint8_t flags = S.flags[i];
        active = flags & 1;
        born = (flags >> 1) & 1;
        birth = S.birth ? S.birth[i] : -INFINITY;
        death = S.death[i];
        arrival = S.arrival[i];
        nc = S.ncoll[i];
    }
    const SpeciesDev sp = P.species[sidx];
    const double inv_m = 1.0 / m;
    uint32_t pushed = 0, ncoll_launch = 0;
    bool checked = false;
    const double dt = P.dt;
    const double half_dt = dt * 0.5;
    __shared__ double4 s_atoms[4][kWaveAtoms];   // 256 threads = 4 wavefronts
    double4* const lds = s_atoms[threadIdx.x >> 6];
    // the next step's table entry is loaded one step ahead
    StepEntry te_next = P.n_steps ? P.table[0] : StepEntry{};
    for (uint32_t k = 0; k < P.n_steps; ++k) {
        // unborn ions wait for their birth, live ones step; none left: the
        // whole wavefront leaves together
        if (__ballot(have && (!born || active)) == 0ull) break;
        const StepEntry te = te_next;
        if (k + 1 < P.n_steps) te_next = P.table[k + 1];
        const double t = te.t;
        bool live = false;
        if (have) {
            if (!born && t >= birth) {
                born = true;
                active = true;
                checked = false;
            }
            if (born && active) {
                if (!checked && !contains_f64(P, pos)) {
                    active = false;
                    death = t;
                } else {
                    live = true;
                }
            }
        }
        const double z_before = pos.z;
        if (live) {
            if constexpr (INTEG == 0) {
                const d3 p0 = pos, v0 = vel;
                const d3 k1v = v0;
                const d3 k1a = force_f64<FIELD, COLL_EHSS>(P, p0, v0, te.c0, te.e0, q, m, 0.0) * inv_m;
                const d3 k2v = v0 + k1a * half_dt;
                const d3 k2a = force_f64<FIELD, COLL_EHSS>(P, p0 + k1v * half_dt, k2v, te.ch, te.eh, q, m, 0.0) * inv_m;
                const d3 k3v = v0 + k2a * half_dt;
                const d3 k3a = force_f64<FIELD, COLL_EHSS>(P, p0 + k2v * half_dt, k3v, te.ch, te.eh, q, m, 0.0) * inv_m;
                const d3 k4v = v0 + k3a * dt;
                const d3 k4a = force_f64<FIELD, COLL_EHSS>(P, p0 + k3v * dt, k4v, te.c1, te.e1, q, m, 0.0) * inv_m;
                const double h6 = dt / 6.0;
                pos = p0 + (k1v + k2v * 2.0 + k3v * 2.0 + k4v) * h6;
                vel = v0 + (k1a + k2a * 2.0 + k3a * 2.0 + k4a) * h6;
            } else {
                const d3 a = force_f64<FIELD, COLL_EHSS>(P, pos, vel, te.c0, te.e0, q, m, 0.0) * inv_m;
                const d3 vm = vel + a * half_dt;
                d3 vp = vm;
                if (P.magnetic) {
                    const double qm = q * inv_m;
                    const d3 tv = d3{P.B[0], P.B[1], P.B[2]} * (qm * dt * 0.5);
                    const double s_f = 2.0 / (1.0 + dot(tv, tv));
                    const d3 sv = tv * s_f;
                    const d3 vprime = vm + cross(vm, tv);
                    vp = vm + cross(vprime, sv);
                }
                pos = pos + vp * dt;
                vel = vp + a * half_dt;
            }
            ++pushed;
        }
        // the collision test of each lane (collide_f64<COLL_EHSS> up to the
        // geometry method)
        PhiloxStream rng(P.seed, i, P.step0 + k);
        bool hit = false;
        d3 vn{0.0, 0.0, 0.0};
        if (live && sp.sigma_ehss > 0.0) {
            double u1 = rng.uniform_f64(), u2 = rng.uniform_f64();
            vn.x = P.sigma_neutral * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
            u1 = rng.uniform_f64(); u2 = rng.uniform_f64();
            vn.y = P.sigma_neutral * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
            u1 = rng.uniform_f64(); u2 = rng.uniform_f64();
            vn.z = P.sigma_neutral * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
            vn = vn + d3{P.u_gas[0], P.u_gas[1], P.u_gas[2]};
            const d3 vrel = vel - vn;
            const double vrel_mag = sqrt(dot(vrel, vrel));
            if (vrel_mag >= kMinRelativeVelocity) {
                const double Pc = 1.0 - exp(-P.n * sp.sigma_ehss * vrel_mag * (dt * te.nrel));
                hit = rng.uniform_f64() < Pc;
            }
        }
        // the geometry method of each colliding ion, on the whole wavefront
        uint64_t pending = __ballot(hit);
        while (pending) {
            const int L = __builtin_amdgcn_readfirstlane(static_cast<int>(__builtin_ctzll(pending)));
            pending &= pending - 1ull;
            WaveStream r;
            r.k0 = rng.k0;
            r.k1 = rng.k1;
            r.s_lo = rng.s_lo;
            r.s_hi = rng.s_hi;
            r.ion = readlane_u32(i, L);
            r.block = readlane_u32(rng.block, L);
#pragma unroll
            for (int c = 0; c < 4; ++c) r.buf[c] = readlane_u32(rng.buf[c], L);
            r.pos = static_cast<int>(readlane_u32(static_cast<uint32_t>(rng.pos), L));
            const d3 v_new = ehss_wave_f64(P, readlane_d3(vel, L), readlane_d3(vn, L), readlane_f64(m, L),
                                           P.m_gas, P.r_gas, readlane_u32(sp.atom_offset, L),
                                           readlane_u32(sp.atom_count, L), r, lane, lds);
            if (static_cast<int>(lane) == L) {
                vel = v_new;
                ++nc;
                ++ncoll_launch;
            }
        }
        if (live) {
            if (P.detect && isnan(arrival)) {
                if (z_before < P.det_z && pos.z >= P.det_z)
                    arrival = t + dt * (P.det_z - z_before) / (pos.z - z_before);
            }
            if (!contains_f64(P, pos)) {
                active = false;
                death = t;
            } else {
                checked = true;
                if (!(isfinite(pos.x) && isfinite(pos.y) && isfinite(pos.z) && isfinite(vel.x) &&
                      isfinite(vel.y) && isfinite(vel.z))) {
                    active = false;
                    death = t;
                }
            }
        }
    }
    if (have) {
        S.x[i] = pos.x; S.y[i] = pos.y; S.z[i] = pos.z;
        S.vx[i] = vel.x; S.vy[i] = vel.y; S.vz[i] = vel.z;
        S.flags[i] = static_cast<uint8_t>((active ? 1 : 0) | (born ? 2 : 0));
        S.death[i] = death;
        S.arrival[i] = arrival;
        S.ncoll[i] = nc;
    }
    add_counters(P, pushed, ncoll_launch);
}

// ======================= FP32 + Kahan (fast path) ===========================

// the FP32 helpers (fields, containment, collisions) are in step_device_f32.hpp

template <int INTEG, int FIELD, int COLL, bool SUB = false, bool FLOW = false>
__device__ __forceinline__ void f32_ion(const StepParams& P, const IonPtrs& S, const uint32_t i,
        uint32_t& pushed, uint32_t& ncoll_launch) {
    const FastConst& C = P.fc;
    // state: FP32 (hi, lo) pairs; hi + lo equals the FP64 value to ~2^-48
    const double X[6] = {S.x[i], S.y[i], S.z[i], S.vx[i], S.vy[i], S.vz[i]};
    float hi[6], lo[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        hi[c] = (float)X[c];
        lo[c] = (float)(X[c] - (double)hi[c]);
    }
    const double md = S.mass[i];
    const double qd = S.charge[i];
    const uint32_t sidx = S.species[i];
    const SpeciesDev sp = P.species[sidx];
    const float m = (float)md;
    const float qm = (float)(qd / md);
    const float inv_m_plus = (float)(1.0 / (md + P.m_gas));
    const float kbt2_dt_over_m = (float)(2.0 * kBoltzmann * P.T / md * P.dt);
    uint8_t flags = S.flags[i];
    bool active = flags & 1, born = (flags >> 1) & 1;
    const double birth = S.birth ? S.birth[i] : -INFINITY;
    double death = S.death[i];
    double arrival = S.arrival[i];
    uint32_t nc = S.ncoll[i];
    // damping factor in units of 1/s (force per mass): Langevin lpre/|v|
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
        lpre = (float)l;
    }
    bool checked = false;
    const float dt = (float)P.dt;
    const float half_dt = 0.5f * dt;
    const float h6 = dt * (1.0f / 6.0f);
    for (uint32_t k = 0; k < P.n_steps; ++k) {
        const StepEntry te = P.table[k];
        const double t = te.t;
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
        if (!checked && !contains_f32(P, C, hi[0], hi[1], hi[2])) {
            active = false;
            death = t;
            break;
        }
        const float z_before = hi[2];
        const float z_before_lo = lo[2];
        // the step's gas density over the domain's (scales gamma and the rates)
        const float nrel = (float)te.nrel;
        const float lp = lpre * nrel;
        const f3 p0{hi[0], hi[1], hi[2]};
        const f3 v0{hi[3], hi[4], hi[5]};
        f3 dpos, dvel;
        if constexpr (INTEG == 0) {
            const f3 k1a = accel_f32<FIELD, COLL, FLOW>(P, C, p0, v0, te.c0, te.e0, qm, lp);
            const f3 k2v = v0 + k1a * half_dt;
            const f3 k2a = accel_f32<FIELD, COLL, FLOW>(P, C, p0 + v0 * half_dt, k2v, te.ch, te.eh, qm, lp);
            const f3 k3v = v0 + k2a * half_dt;
            const f3 k3a = accel_f32<FIELD, COLL, FLOW>(P, C, p0 + k2v * half_dt, k3v, te.ch, te.eh, qm, lp);
            const f3 k4v = v0 + k3a * dt;
            const f3 k4a = accel_f32<FIELD, COLL, FLOW>(P, C, p0 + k3v * dt, k4v, te.c1, te.e1, qm, lp);
            dpos = (v0 + k2v * 2.0f + k3v * 2.0f + k4v) * h6;
            dvel = (k1a + k2a * 2.0f + k3a * 2.0f + k4a) * h6;
        } else {
            const f3 a = accel_f32<FIELD, COLL, FLOW>(P, C, p0, v0, te.c0, te.e0, qm, lp);
            const f3 vm = v0 + a * half_dt;
            f3 rot{0.0f, 0.0f, 0.0f};
            if (P.magnetic) {
                const f3 tv = f3{(float)P.B[0], (float)P.B[1], (float)P.B[2]} * (qm * half_dt);
                const float s_f = 2.0f / (1.0f + dot(tv, tv));
                const f3 vprime = vm + cross(vm, tv);
                rot = cross(vprime, tv * s_f);
            }
            // x += v+ dt, v = v+ + a dt/2 = v0 + a dt + rot (no cancellation)
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
            PhiloxStream rng(P.seed, i, P.step0 + k);
            f3 v{hi[3], hi[4], hi[5]};
            const f3 vin = v;
            // collision substeps: dt / n_sub each, one continuing stream, in the
            // SUB instantiation (the loop costs the one-stage kernels a
            // wavefront per SIMD)
            const int ns = SUB ? P.n_sub : 1;
            const float sub = SUB ? 1.0f / static_cast<float>(P.n_sub) : 1.0f;
            const float dtc = SUB ? static_cast<float>(P.dt / static_cast<double>(P.n_sub)) : dt;
            uint32_t hits = 0;
            const f3 ug = FLOW ? gas_velocity_f32(P, C, f3{hi[0], hi[1], hi[2]}) : f3{0.0f, 0.0f, 0.0f};
            for (int s = 0; s < ns; ++s)
                hits += collide_f32<COLL, FLOW>(P, C, sp, sidx, m, inv_m_plus, lp, kbt2_dt_over_m * sub, v, rng,
                                                dtc * nrel, C.n_dt * sub * nrel, ug) ? 1u : 0u;
            if (COLL == COLL_HSS || COLL == COLL_HSS_MIX) {
                if (hits) {
                    // a collision replaces the velocity: restart its compensation
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
            if (zb < P.det_z && za >= P.det_z)
                arrival = t + P.dt * (P.det_z - zb) / (za - zb);
        }
        if (!contains_f32(P, C, hi[0], hi[1], hi[2])) {
            active = false;
            death = t;
            break;
        }
        checked = true;
        if (!(isfinite(hi[0]) && isfinite(hi[1]) && isfinite(hi[2]) && isfinite(hi[3]) &&
              isfinite(hi[4]) && isfinite(hi[5]))) {
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
}

template <int INTEG, int FIELD, int COLL, bool SUB>
__global__ void __launch_bounds__(256) step_f32_kernel(const StepParams P, const IonPtrs S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t pushed = 0, ncoll_launch = 0;
    if (i < P.n_ions) f32_ion<INTEG, FIELD, COLL, SUB, SUB>(P, S, P.order ? P.order[i] : i, pushed, ncoll_launch);
    add_counters(P, pushed, ncoll_launch);
}

// ======================= FP32 + EHSS surrogate ==============================
//
// EHSS with the scattering normal drawn by the surrogate MLP instead of the
// geometry ray test.  The collision test is the reference's (neutral from the
// Maxwellian, P = 1 - exp(-n sigma_eff |v_rel| dt), EHSSCollisionHandler.cpp:222-246);
// a colliding ion draws (u, psi), the MLP maps (u, cos psi, sin psi, features)
// to the contact normal in the collision frame, and the relative velocity is
// reflected about it (offline-table branch, EHSSCollisionHandler.cpp:253-277).
// mlp_eval(step, coll, in, out) evaluates the MLP for the colliding lanes:
// mlp_group_eval (the single-cloud kernel: the workgroup's colliding ions
// gathered into 64-ion batches through LDS, so the matrix cores evaluate about
// as many rows as there are collisions) or mlp_wave_eval_any (the batch
// kernel: per wavefront, no barriers).  Every lane stays in the step loop
// (dead lanes are masked, not exited) so that barriers and the wave-wide matrix
// instructions see every lane they need.  One ion per lane over P.n_steps
// steps; with mlp_group_eval every lane of the workgroup calls it with the same
// P.n_steps (lanes with i >= P.n_ions are carried along masked).
template <int INTEG, int FIELD, class MlpEval>
__device__ __forceinline__ void surr_ion(const StepParams& P, const IonPtrs& S, uint32_t i,
                                         MlpEval&& mlp_eval, uint32_t& pushed, uint32_t& ncoll_launch) {
    const bool valid = i < P.n_ions;
    const uint32_t ii = valid ? i : 0;
    const FastConst& C = P.fc;
    const double X[6] = {S.x[ii], S.y[ii], S.z[ii], S.vx[ii], S.vy[ii], S.vz[ii]};
    float hi[6], lo[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        hi[c] = (float)X[c];
        lo[c] = (float)(X[c] - (double)hi[c]);
    }
    const double md = S.mass[ii];
    const double qd = S.charge[ii];
    const SpeciesDev sp = P.species[S.species[ii]];
    const float sigma = (float)sp.sigma_ehss;
    const float qm = (float)(qd / md);
    const float m = (float)md;
    const float inv_m_plus = (float)(1.0 / (md + P.m_gas));
    uint8_t flags = S.flags[ii];
    bool active = valid && (flags & 1), born = valid && ((flags >> 1) & 1);
    const double birth = S.birth ? S.birth[ii] : -INFINITY;
    double death = S.death[ii];
    double arrival = S.arrival[ii];
    uint32_t nc = S.ncoll[ii];
    bool checked = false;
    const float dt = (float)P.dt;
    const float half_dt = 0.5f * dt;
    const float h6 = dt * (1.0f / 6.0f);
    // the next step's table entry is loaded one step ahead
    StepEntry te_next = P.n_steps ? P.table[0] : StepEntry{};
    for (uint32_t k = 0; k < P.n_steps; ++k) {
        const StepEntry te = te_next;
        if (k + 1 < P.n_steps) te_next = P.table[k + 1];
        const double t = te.t;
        if (valid && !born && t >= birth) {
            born = true;
            active = true;
            checked = false;
        }
        bool live = active && born;
        if (live && !checked && !contains_f32(P, C, hi[0], hi[1], hi[2])) {
            active = false;
            death = t;
            live = false;
        }
        bool coll = false;
        float in[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        f3 vn{0.f, 0.f, 0.f}, vrel{0.f, 0.f, 0.f};
        float vm = 0.f;
        float z_before = hi[2], z_before_lo = lo[2];
        if (live) {
            const f3 p0{hi[0], hi[1], hi[2]};
            const f3 v0{hi[3], hi[4], hi[5]};
            f3 dpos, dvel;
            if constexpr (INTEG == 0) {
                const f3 k1a = accel_f32<FIELD, COLL_NONE>(P, C, p0, v0, te.c0, te.e0, qm, 0.f);
                const f3 k2v = v0 + k1a * half_dt;
                const f3 k2a = accel_f32<FIELD, COLL_NONE>(P, C, p0 + v0 * half_dt, k2v, te.ch, te.eh, qm, 0.f);
                const f3 k3v = v0 + k2a * half_dt;
                const f3 k3a = accel_f32<FIELD, COLL_NONE>(P, C, p0 + k2v * half_dt, k3v, te.ch, te.eh, qm, 0.f);
                const f3 k4v = v0 + k3a * dt;
                const f3 k4a = accel_f32<FIELD, COLL_NONE>(P, C, p0 + k3v * dt, k4v, te.c1, te.e1, qm, 0.f);
                dpos = (v0 + k2v * 2.0f + k3v * 2.0f + k4v) * h6;
                dvel = (k1a + k2a * 2.0f + k3a * 2.0f + k4a) * h6;
            } else {
                const f3 a = accel_f32<FIELD, COLL_NONE>(P, C, p0, v0, te.c0, te.e0, qm, 0.f);
                const f3 vmid = v0 + a * half_dt;
                f3 rot{0.0f, 0.0f, 0.0f};
                if (P.magnetic) {
                    const f3 tv = f3{(float)P.B[0], (float)P.B[1], (float)P.B[2]} * (qm * half_dt);
                    const float s_f = 2.0f / (1.0f + dot(tv, tv));
                    const f3 vprime = vmid + cross(vmid, tv);
                    rot = cross(vprime, tv * s_f);
                }
                dpos = (vmid + rot) * dt;
                dvel = a * dt + rot;
            }
            kahan_add(hi[0], lo[0], dpos.x);
            kahan_add(hi[1], lo[1], dpos.y);
            kahan_add(hi[2], lo[2], dpos.z);
            kahan_add(hi[3], lo[3], dvel.x);
            kahan_add(hi[4], lo[4], dvel.y);
            kahan_add(hi[5], lo[5], dvel.z);
            ++pushed;
            if (sigma > 0.0f) {
                PhiloxStream rng(P.seed, i, P.step0 + k);
                const float u1 = rng.uniform_f32(), u2 = rng.uniform_f32();
                const float u3 = rng.uniform_f32(), u4 = rng.uniform_f32();
                const float r1 = C.sigma_neutral * fast_sqrt(-2.0f * fast_ln(u1));
                const float r2 = C.sigma_neutral * fast_sqrt(-2.0f * fast_ln(u3));
                vn = f3{r1 * fast_cos_rev(u2) + C.u_gas[0], r1 * fast_sin_rev(u2) + C.u_gas[1],
                        r2 * fast_cos_rev(u4) + C.u_gas[2]};
                const f3 v{hi[3], hi[4], hi[5]};
                vrel = v - vn;
                vm = fast_sqrt(dot(vrel, vrel));
                if (vm >= (float)kMinRelativeVelocity) {
                    // n dt of the step (the density ratio of a pressure waveform)
                    const float x = C.n_dt * (float)te.nrel * sigma * vm;
                    const float Pc = x < 1e-3f ? x * (1.0f - 0.5f * x * (1.0f - x * (1.0f / 3.0f)))
                                               : 1.0f - __builtin_amdgcn_exp2f(-x * 1.44269504089f);
                    if (rng.uniform_f32() < Pc) {
                        coll = true;
                        const float uq = rng.uniform_f32();
                        const float psi = rng.uniform_f32();  // revolutions
                        in[0] = uq;
                        in[1] = fast_cos_rev(psi);
                        in[2] = fast_sin_rev(psi);
                        in[3] = sp.feat[0];
                        in[4] = sp.feat[1];
                        in[5] = sp.feat[2];
                    }
                }
            }
        }
        float out[3];
        if (mlp_eval(k, coll, in, out)) {
            if (coll) {
                const float nrm = fast_sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
                float mu = nrm > 0.f ? -out[2] / nrm : 0.f;
                mu = fminf(1.f, fmaxf(0.f, mu));
                const float st = fast_sqrt(fmaxf(0.f, 1.f - mu * mu));
                const float inv = 1.0f / vm;
                const f3 ax{vrel.x * inv, vrel.y * inv, vrel.z * inv};
                const f3 a = fabsf(ax.x) < 0.9f ? f3{1.f, 0.f, 0.f} : f3{0.f, 1.f, 0.f};
                f3 t1 = cross(ax, a);
                t1 = t1 * (1.0f / fast_sqrt(dot(t1, t1)));
                f3 t2 = cross(ax, t1);
                t2 = t2 * (1.0f / fast_sqrt(dot(t2, t2)));
                const f3 nvec = t1 * (st * in[1]) + t2 * (st * in[2]) + ax * (-mu);
                const f3 vr = vrel - nvec * (2.0f * dot(vrel, nvec));
                const f3 v{hi[3], hi[4], hi[5]};
                const f3 Vcom = (v * m + vn * C.m_gas) * inv_m_plus;
                const f3 vnew = Vcom + vr * (C.m_gas * inv_m_plus);
                hi[3] = vnew.x; hi[4] = vnew.y; hi[5] = vnew.z;
                lo[3] = 0.f; lo[4] = 0.f; lo[5] = 0.f;
                ++nc;
                ++ncoll_launch;
            }
        }
        if (live) {
            if (P.detect && isnan(arrival)) {
                const double zb = (double)z_before + (double)z_before_lo;
                const double za = (double)hi[2] + (double)lo[2];
                if (zb < P.det_z && za >= P.det_z) arrival = t + P.dt * (P.det_z - zb) / (za - zb);
            }
            if (!contains_f32(P, C, hi[0], hi[1], hi[2]) ||
                !(isfinite(hi[0]) && isfinite(hi[1]) && isfinite(hi[2]) && isfinite(hi[3]) &&
                  isfinite(hi[4]) && isfinite(hi[5]))) {
                active = false;
                death = t;
            } else {
                checked = true;
            }
        }
    }
    if (valid) {
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
    }
}

template <int INTEG, int FIELD>
__global__ void __launch_bounds__(kSurrogateBlock, 3) step_surrogate_kernel(const StepParams P, const IonPtrs S) {
    __shared__ uint4 mlp_lds[kPackBytes / 16];
    __shared__ MlpExchange<kSurrogateBlock / 64> xch;
    mlp_exchange_init(xch);
    mlp_stage(P.mlp_blob, mlp_lds);
    const unsigned char* lds = reinterpret_cast<const unsigned char*>(mlp_lds);
    uint32_t pushed = 0, nc = 0;
    surr_ion<INTEG, FIELD>(P, S, blockIdx.x * blockDim.x + threadIdx.x,
                           [&](uint32_t k, bool coll, const float* in, float* out) {
                               return mlp_group_eval(lds, xch, k, coll, in, out, k + blockIdx.x);
                           },
                           pushed, nc);
    add_counters(P, pushed, nc);
}

// ======================= batched clouds (persistent) ========================
//
// A fixed pool of workgroups pulls tiles from a queue until it is empty; the
// host orders the tiles longest first, so the long clouds start first and the
// short ones fill in at the end.  Every workgroup leaves the loop once the
// queue index passes the tile count.  Each tile runs the same per-ion body as
// the single-cloud kernels, so a cloud's result does not depend on the batch.

template <int PREC, bool SUB, int I, int F, int Cc>
__device__ __forceinline__ void batch_ion(const StepParams& P, const IonPtrs& S, uint32_t i,
                                          uint32_t& pushed, uint32_t& nc) {
    if constexpr (PREC == 1) {
        f32_ion<I, F, Cc, SUB>(P, S, i, pushed, nc);
    } else {
        f64_ion<I, F, Cc, SUB>(P, S, i, pushed, nc);
    }
}

// MIX: the kernel also holds the gas-mixture kinds (12-15), in FP64 the
// geometry EHSS kinds (16-19), and the uniform-grid field-array kinds (20-25),
// and its FP32 bodies run collision substeps (clouds with more than one
// collision stage per step run in this variant).
// Their bodies need more registers, so a batch without such clouds runs the
// variant without them (FP32: 90 VGPRs, 5 waves per SIMD).
template <int PREC, bool MIX>
__device__ __forceinline__ void batch_dispatch(uint32_t kind, const StepParams& P, const IonPtrs& S,
                                               uint32_t i, uint32_t& pushed, uint32_t& nc) {
    switch (kind) {
        case 0: batch_ion<PREC, MIX, 0, FIELD_LINEAR, COLL_NONE>(P, S, i, pushed, nc); break;
        case 1: batch_ion<PREC, MIX, 1, FIELD_LINEAR, COLL_NONE>(P, S, i, pushed, nc); break;
        case 2: batch_ion<PREC, MIX, 0, FIELD_ORBITRAP, COLL_NONE>(P, S, i, pushed, nc); break;
        case 3: batch_ion<PREC, MIX, 1, FIELD_ORBITRAP, COLL_NONE>(P, S, i, pushed, nc); break;
        case 4: batch_ion<PREC, MIX, 0, FIELD_LINEAR, COLL_HSS>(P, S, i, pushed, nc); break;
        case 5: batch_ion<PREC, MIX, 1, FIELD_LINEAR, COLL_HSS>(P, S, i, pushed, nc); break;
        case 6: batch_ion<PREC, MIX, 0, FIELD_ORBITRAP, COLL_HSS>(P, S, i, pushed, nc); break;
        case 7: batch_ion<PREC, MIX, 1, FIELD_ORBITRAP, COLL_HSS>(P, S, i, pushed, nc); break;
        case 8: batch_ion<PREC, MIX, 0, FIELD_LINEAR, COLL_DAMPING>(P, S, i, pushed, nc); break;
        case 9: batch_ion<PREC, MIX, 1, FIELD_LINEAR, COLL_DAMPING>(P, S, i, pushed, nc); break;
        case 10: batch_ion<PREC, MIX, 0, FIELD_ORBITRAP, COLL_DAMPING>(P, S, i, pushed, nc); break;
        case 11: batch_ion<PREC, MIX, 1, FIELD_ORBITRAP, COLL_DAMPING>(P, S, i, pushed, nc); break;
        default:
            if constexpr (MIX) {
                switch (kind) {
                    case 12: batch_ion<PREC, MIX, 0, FIELD_LINEAR, COLL_HSS_MIX>(P, S, i, pushed, nc); break;
                    case 13: batch_ion<PREC, MIX, 1, FIELD_LINEAR, COLL_HSS_MIX>(P, S, i, pushed, nc); break;
                    case 14: batch_ion<PREC, MIX, 0, FIELD_ORBITRAP, COLL_HSS_MIX>(P, S, i, pushed, nc); break;
                    case 15: batch_ion<PREC, MIX, 1, FIELD_ORBITRAP, COLL_HSS_MIX>(P, S, i, pushed, nc); break;
                    case 20: batch_ion<PREC, MIX, 0, FIELD_ARRAY, COLL_NONE>(P, S, i, pushed, nc); break;
                    case 21: batch_ion<PREC, MIX, 1, FIELD_ARRAY, COLL_NONE>(P, S, i, pushed, nc); break;
                    case 22: batch_ion<PREC, MIX, 0, FIELD_ARRAY, COLL_HSS>(P, S, i, pushed, nc); break;
                    case 23: batch_ion<PREC, MIX, 1, FIELD_ARRAY, COLL_HSS>(P, S, i, pushed, nc); break;
                    case 24: batch_ion<PREC, MIX, 0, FIELD_ARRAY, COLL_DAMPING>(P, S, i, pushed, nc); break;
                    case 25: batch_ion<PREC, MIX, 1, FIELD_ARRAY, COLL_DAMPING>(P, S, i, pushed, nc); break;
                    default:
                        if constexpr (PREC == 0) {
                            switch (kind) {
                                case 16: batch_ion<0, MIX, 0, FIELD_LINEAR, COLL_EHSS>(P, S, i, pushed, nc); break;
                                case 17: batch_ion<0, MIX, 1, FIELD_LINEAR, COLL_EHSS>(P, S, i, pushed, nc); break;
                                case 18: batch_ion<0, MIX, 0, FIELD_ORBITRAP, COLL_EHSS>(P, S, i, pushed, nc); break;
                                case 19: batch_ion<0, MIX, 1, FIELD_ORBITRAP, COLL_EHSS>(P, S, i, pushed, nc); break;
                                default: break;
                            }
                        }
                        break;
                }
            }
            break;
    }
}

// Persistent tile loop shared by the batch kernels: take tiles from the queue
// until it is empty, run `body(cloud, first ion)` on each, stamp the cloud's
// completion clock after the tile.  Queue and clock updates are made by wave 0
// under a wave-uniform branch (all its lanes, one combined increment): a
// lane-divergent region on each side of the loop's back edge could be merged
// by the compiler into a divergent loop in which lanes 1..63 pass the barriers
// again before lane 0 has taken the next tile.
template <typename Body>
__device__ __forceinline__ void batch_queue_loop(const BatchTile* __restrict__ tiles, uint32_t n_tiles,
                                                 uint32_t* queue, unsigned long long* done,
                                                 unsigned long long* start, Body&& body) {
    __shared__ uint32_t s_tile;
    const bool wave0 = __builtin_amdgcn_readfirstlane(threadIdx.x) == 0;
    const uint32_t lane = threadIdx.x & 63;
    if (wave0) atomicMin(start, static_cast<unsigned long long>(__builtin_amdgcn_s_memrealtime()));
    uint32_t prev = 0xFFFFFFFFu;
    for (;;) {
        if (wave0) {
            if (prev != 0xFFFFFFFFu)
                atomicMax(&done[prev], static_cast<unsigned long long>(__builtin_amdgcn_s_memrealtime()));
            const uint32_t got = atomicAdd(queue, lane == 0 ? 1u : 0u);
            s_tile = __builtin_amdgcn_readfirstlane(got);
        }
        __syncthreads();
        const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();
        if (tile >= n_tiles) break;
        const uint32_t c = __builtin_amdgcn_readfirstlane(tiles[tile].cloud);
        const uint32_t first = __builtin_amdgcn_readfirstlane(tiles[tile].first);
        body(c, first);
        __syncthreads();   // the whole tile is done before its cloud's clock is stamped
        prev = c;
    }
}

// The per-cloud tables come in as separate read-only, non-aliased arguments so
// that the wave-uniform parameter loads stay scalar.
template <int PREC, bool MIX>
__global__ void __launch_bounds__(256) batch_kernel(const StepParams* __restrict__ params,
                                                    const IonPtrs* __restrict__ ions,
                                                    const uint32_t* __restrict__ kinds,
                                                    const BatchTile* __restrict__ tiles, uint32_t n_tiles,
                                                    uint32_t* queue, unsigned long long* done,
                                                    unsigned long long* start) {
    batch_queue_loop(tiles, n_tiles, queue, done, start, [&](uint32_t c, uint32_t first) {
        const StepParams& P = params[c];
        const IonPtrs S = ions[c];
        const uint32_t kind = __builtin_amdgcn_readfirstlane(kinds[c]);
        const uint32_t i = first + threadIdx.x;
        uint32_t pushed = 0, nc = 0;
        if (i < P.n_ions) batch_dispatch<PREC, MIX>(kind, P, S, i, pushed, nc);
        add_counters(P, pushed, nc);
    });
}

// FP32 + EHSS surrogate clouds: the packed MLP weights are staged into LDS once
// per persistent workgroup and serve every tile it takes.  kind = integrator
// (0 RK4, 1 Boris) + 2 * (Orbitrap field).  The MLP runs per wavefront, both
// halves together (mlp_wave_eval_any, two waves per SIMD): in this kernel the
// workgroup-wide batches and the one-block-at-a-time core of the single-cloud
// kernel measured slower (2.53e10 against 2.88e10 ion-steps/s on 256 EHSS
// clouds).  A wavefront whose 64 ions all lie beyond the cloud's end skips the
// tile.
__global__ void __launch_bounds__(256, 2) batch_surrogate_kernel(const StepParams* __restrict__ params,
                                                                 const IonPtrs* __restrict__ ions,
                                                                 const uint32_t* __restrict__ kinds,
                                                                 const BatchTile* __restrict__ tiles,
                                                                 uint32_t n_tiles, uint32_t* queue,
                                                                 unsigned long long* done,
                                                                 unsigned long long* start,
                                                                 const uint4* __restrict__ blob) {
    __shared__ uint4 mlp_lds[kPackBytes / 16];
    mlp_stage(blob, mlp_lds);
    const unsigned char* lds = reinterpret_cast<const unsigned char*>(mlp_lds);
    const auto eval = [&](uint32_t, bool coll, const float* in, float* out) {
        return mlp_wave_eval_any(lds, coll, in, out);
    };
    batch_queue_loop(tiles, n_tiles, queue, done, start, [&](uint32_t c, uint32_t first) {
        const StepParams& P = params[c];
        const IonPtrs S = ions[c];
        const uint32_t kind = __builtin_amdgcn_readfirstlane(kinds[c]);
        const uint32_t wave_first = first + __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
        uint32_t pushed = 0, nc = 0;
        if (wave_first < P.n_ions) {
            const uint32_t i = first + threadIdx.x;
            switch (kind) {
                case 0: surr_ion<0, FIELD_LINEAR>(P, S, i, eval, pushed, nc); break;
                case 1: surr_ion<1, FIELD_LINEAR>(P, S, i, eval, pushed, nc); break;
                case 2: surr_ion<0, FIELD_ORBITRAP>(P, S, i, eval, pushed, nc); break;
                case 3: surr_ion<1, FIELD_ORBITRAP>(P, S, i, eval, pushed, nc); break;
                default: break;
            }
        }
        add_counters(P, pushed, nc);
    });
}

__global__ void batch_reset_kernel(uint32_t* queue, unsigned long long* done, unsigned long long* start,
                                   unsigned long long* counters, uint32_t n_clouds) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kBatchClasses) queue[i] = 0u;
    if (i == 0) *start = ~0ull;
    if (i < n_clouds) {
        done[i] = 0ull;
        counters[4 * i + 0] = 0ull; counters[4 * i + 1] = 0ull;
        counters[4 * i + 2] = 0ull; counters[4 * i + 3] = 0ull;
    }
}

// Reset the queue/clock/counter records, then the persistent kernel of each
// class present (FP64, FP32, FP32 + surrogate; a class with n_tiles == 0 is
// skipped).  The first class present runs on `s`, the others on the side
// streams, forked from and joined back into `s`, so that the classes overlap.
hipError_t launch_batch(const BatchParams* B, const uint32_t* block, const uint32_t* grid,
                        const uint4* blob, uint32_t n_clouds, unsigned long long* counters, hipStream_t s,
                        const hipStream_t* side, hipEvent_t fork, const hipEvent_t* join) {
    hipLaunchKernelGGL(batch_reset_kernel, dim3((n_clouds + 255) / 256 + 1), dim3(256), 0, s, B[0].queue,
                       B[0].done, B[0].start, counters, n_clouds);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int present[kBatchClasses], n_present = 0;
    for (int c = 0; c < kBatchClasses; ++c)
        if (B[c].n_tiles) present[n_present++] = c;
    if (n_present > 1 && (e = hipEventRecord(fork, s)) != hipSuccess) return e;
    for (int k = 0; k < n_present; ++k) {
        const int c = present[k];
        hipStream_t st = k == 0 ? s : side[k - 1];
        if (k > 0 && (e = hipStreamWaitEvent(st, fork, 0)) != hipSuccess) return e;
        const BatchParams& b = B[c];
        if (c == 0 && b.with_mix)
            hipLaunchKernelGGL((batch_kernel<0, true>), dim3(grid[c]), dim3(block[c]), 0, st, b.params, b.ions,
                               b.kinds, b.tiles, b.n_tiles, b.queue, b.done, b.start);
        else if (c == 0)
            hipLaunchKernelGGL((batch_kernel<0, false>), dim3(grid[c]), dim3(block[c]), 0, st, b.params, b.ions,
                               b.kinds, b.tiles, b.n_tiles, b.queue, b.done, b.start);
        else if (c == 1 && b.with_mix)
            hipLaunchKernelGGL((batch_kernel<1, true>), dim3(grid[c]), dim3(block[c]), 0, st, b.params, b.ions,
                               b.kinds, b.tiles, b.n_tiles, b.queue, b.done, b.start);
        else if (c == 1)
            hipLaunchKernelGGL((batch_kernel<1, false>), dim3(grid[c]), dim3(block[c]), 0, st, b.params, b.ions,
                               b.kinds, b.tiles, b.n_tiles, b.queue, b.done, b.start);
        else
            hipLaunchKernelGGL(batch_surrogate_kernel, dim3(grid[c]), dim3(256), 0, st, b.params, b.ions,
                               b.kinds, b.tiles, b.n_tiles, b.queue, b.done, b.start, blob);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    for (int k = 1; k < n_present; ++k) {
        if ((e = hipEventRecord(join[k - 1], side[k - 1])) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(s, join[k - 1], 0)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// ======================= trajectory snapshots ================================
//
// One pass over the SoA state per snapshot:
//   * the full snapshot in OutputManager's layout (positions and velocities
//     interleaved [n][3] in FP64, species, domain index, flags) into a
//     device staging slot;
//   * the records of the ions that may have changed since the previous
//     snapshot: those active now, or whose flags changed (died or were born
//     since).  Unborn and dead ions do not move, so the others hold the values
//     the previous snapshot gave them.  Records are compacted per workgroup
//     (wave ballots + one atomic per workgroup); their order across workgroups
//     is arbitrary and each carries its ion index.
// prev[] keeps every ion's flags at the last snapshot; full = 1 selects every ion.
struct SnapRecord {
    double p[3];
    double v[3];
    uint32_t idx;
    uint8_t flags;
    int8_t domain;
    uint16_t pad;
};
static_assert(sizeof(SnapRecord) == 56, "snapshot record layout");

// 1024 threads per workgroup: the records of its 16 waves take one atomic.  No
// thread leaves before the barriers (lanes past n carry sel = false).
constexpr int kPackThreads = 1024;

__global__ void __launch_bounds__(kPackThreads) pack_snapshot_kernel(const IonPtrs S, uint32_t n, int full,
                                                                     double* __restrict__ pos, double* __restrict__ vel,
                                                                     uint32_t* __restrict__ spec,
                                                                     int32_t* __restrict__ dom,
                                                                     uint8_t* __restrict__ flags,
                                                                     uint8_t* __restrict__ prev,
                                                                     SnapRecord* __restrict__ rec,
                                                                     uint32_t* __restrict__ count) {
    __shared__ uint32_t wave_base[kPackThreads / 64];
    __shared__ uint32_t block_base;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < n;
    bool sel = false;
    double x = 0, y = 0, z = 0, vx = 0, vy = 0, vz = 0;
    uint8_t f = 0;
    int32_t d = 0;
    if (in) {
        x = S.x[i]; y = S.y[i]; z = S.z[i];
        vx = S.vx[i]; vy = S.vy[i]; vz = S.vz[i];
        f = S.flags[i];
        d = S.domain ? S.domain[i] : 0;
        const size_t o = 3ull * i;
        pos[o] = x; pos[o + 1] = y; pos[o + 2] = z;
        vel[o] = vx; vel[o + 1] = vy; vel[o + 2] = vz;
        spec[i] = S.species[i];
        dom[i] = d;
        flags[i] = f;
        sel = full || (f & 1) || f != prev[i];
        prev[i] = f;
    }
    const unsigned long long m = __ballot(sel);
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int wave = static_cast<int>(threadIdx.x >> 6);
    if (lane == 0) wave_base[wave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        const int n_waves = static_cast<int>(blockDim.x >> 6);
        for (int w = 0; w < n_waves; ++w) {
            const uint32_t c = wave_base[w];
            wave_base[w] = total;
            total += c;
        }
        block_base = total ? atomicAdd(count, total) : 0u;
    }
    __syncthreads();
    if (sel) {
        const uint32_t k = block_base + wave_base[wave] +
                           static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
        SnapRecord r;
        r.p[0] = x; r.p[1] = y; r.p[2] = z;
        r.v[0] = vx; r.v[1] = vy; r.v[2] = vz;
        r.idx = i;
        r.flags = f;
        r.domain = static_cast<int8_t>(d);
        r.pad = 0;
        rec[k] = r;
    }
}

// After a pack: its record count to pinned host memory (a vector store, read by
// the host once the stream has passed this kernel), and the counter of the
// next pack cleared.
__global__ void publish_count_kernel(const uint32_t* count, uint32_t* host_count, uint32_t* next_count) {
    if (threadIdx.x == 0) {
        host_count[0] = count[0];
        next_count[0] = 0u;
        __threadfence_system();
    }
}

__global__ void zero_u32_kernel(uint32_t* p) {
    if (threadIdx.x == 0) p[0] = 0u;
}

hipError_t launch_zero_u32(uint32_t* p, hipStream_t s) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, p);
    return hipGetLastError();
}

// count must be zero on entry (launch_zero_u32 before the first pack; each
// publish clears next_count for the pack after it).
hipError_t launch_pack_snapshot(const IonPtrs& S, uint32_t n, int full, double* pos, double* vel, uint32_t* spec,
                                int32_t* dom, uint8_t* flags, uint8_t* prev, void* rec, uint32_t* count,
                                uint32_t* host_count, uint32_t* next_count, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pack_snapshot_kernel, dim3((n + kPackThreads - 1) / kPackThreads), dim3(kPackThreads), 0, s,
                       S, n, full, pos, vel, spec, dom, flags, prev, static_cast<SnapRecord*>(rec), count);
    hipLaunchKernelGGL(publish_count_kernel, dim3(1), dim3(64), 0, s, count, host_count, next_count);
    return hipGetLastError();
}

// ---- launch dispatch --------------------------------------------------------

using KernelFn = void (*)(const StepParams, const IonPtrs);

// sub: more than one collision stage per step (the FP32 kernels have an
// instantiation of their own for it; the FP64 bodies loop at run time)
template <int I, int F, int Cc>
KernelFn pick_prec(int precision, bool sub) {
    if constexpr (Cc != COLL_NONE) {
        if (sub) return precision == 1 ? step_f32_kernel<I, F, Cc, true> : step_f64_kernel<I, F, Cc, true>;
    }
    return precision == 1 ? step_f32_kernel<I, F, Cc, false> : step_f64_kernel<I, F, Cc, false>;
}

template <int I, int F, int Cc>
KernelFn pick_f64(bool sub) {
    return sub ? step_f64_kernel<I, F, Cc, true> : step_f64_kernel<I, F, Cc, false>;
}

template <int I, int F>
KernelFn pick_coll(int coll, int precision, bool sub) {
    switch (coll) {
        case COLL_HSS: return pick_prec<I, F, COLL_HSS>(precision, sub);
        case COLL_HSS_MIX: return pick_prec<I, F, COLL_HSS_MIX>(precision, sub);
        case COLL_EHSS: return precision == 1 ? nullptr : step_f64_ehss_kernel<I, F>;
        case COLL_EHSS_TAB: return precision == 1 ? nullptr : pick_f64<I, F, COLL_EHSS_TAB>(sub);
        case COLL_EHSS_MIX: return precision == 1 ? nullptr : pick_f64<I, F, COLL_EHSS_MIX>(sub);
        case COLL_DAMPING: return pick_prec<I, F, COLL_DAMPING>(precision, sub);
        default: return pick_prec<I, F, COLL_NONE>(precision, sub);
    }
}

template <int I>
KernelFn pick_field(int field, int coll, int precision, bool sub) {
    switch (field) {
        case FIELD_ORBITRAP: return pick_coll<I, FIELD_ORBITRAP>(coll, precision, sub);
        case FIELD_ARRAY: return pick_coll<I, FIELD_ARRAY>(coll, precision, sub);
        case FIELD_ARRAY_X: return pick_coll<I, FIELD_ARRAY_X>(coll, precision, sub);
        default: return pick_coll<I, FIELD_LINEAR>(coll, precision, sub);
    }
}

template <int I>
KernelFn pick_surrogate(int field) {
    switch (field) {
        case FIELD_ORBITRAP: return step_surrogate_kernel<I, FIELD_ORBITRAP>;
        case FIELD_ARRAY: return step_surrogate_kernel<I, FIELD_ARRAY>;
        case FIELD_ARRAY_X: return step_surrogate_kernel<I, FIELD_ARRAY_X>;
        default: return step_surrogate_kernel<I, FIELD_LINEAR>;
    }
}

// Returns hipSuccess, or hipErrorInvalidValue when the combination has no
// kernel (the host maps that to ICARION_ERR_UNSUPPORTED).
hipError_t launch_step(int integrator, int field, int coll, int precision, int surrogate,
                       const StepParams& P, const IonPtrs& S, uint32_t block, hipStream_t stream) {
    KernelFn fn;
    if (surrogate) {
        fn = integrator == 1 ? pick_surrogate<1>(field) : pick_surrogate<0>(field);
    } else {
        // the extended instantiation: collision substeps or a parabolic flow
        const bool ext = P.n_sub > 1 || P.flow_parabolic;
        fn = integrator == 1 ? pick_field<1>(field, coll, precision, ext) : pick_field<0>(field, coll, precision, ext);
    }
    if (!fn) return hipErrorInvalidValue;
    // the surrogate kernel runs its own workgroup size (the handle's block size
    // does not change any result)
    if (surrogate) block = kSurrogateBlock;
    const uint32_t grid = (P.n_ions + block - 1) / block;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(block), 0, stream, P, S);
    return hipGetLastError();
}

}  // namespace icarion
