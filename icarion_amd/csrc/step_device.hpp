// step_device.hpp - FP64 device building blocks of the engine step shared by
// the fixed-step kernels (step_kernels.hip) and the RK45 kernel
// (rk45_kernels.hip): containment, analytic and array fields, force,
// HSS / EHSS / OU collision stages, wavefront counters.
#pragma once

#include "device_common.hpp"
#include "step_params.hpp"

namespace icarion {

// ======================= FP64 (plain method) ===============================

using d3 = vec3<double>;

__device__ __forceinline__ bool contains_f64(const StepParams& P, d3 p) {
    d3 l{p.x - P.origin[0], p.y - P.origin[1], p.z - P.origin[2]};
    if (!P.rot_identity) l = mat_mul(P.g2l, l);
    if (P.instrument == 3) {
        // OrbitrapGeometry::contains (OrbitrapGeometry.h:36-51).  The electrode
        // radius r_e(z) is the root of f(r) = z^2 - (r^2 - R^2)/2 - Rm^2 ln(R/r),
        // increasing in r below Rm, so r >= r_e(z) <=> f(r) >= 0 (no bisection).
        const double z_abs = fabs(l.z);
        if (z_abs > 0.5 * P.L + kGeomEpsilon) return false;
        const double r2 = l.x * l.x + l.y * l.y;
        const double r = sqrt(r2);
        if (r > P.R_out + kGeomEpsilon) return false;
        const double ln_r = 0.5 * log(r2);
        const double z2 = z_abs * z_abs;
        const double f_in = z2 - 0.5 * (r2 - P.R_in * P.R_in) - P.R_char_sq * (P.ln_R_in - ln_r);
        const double f_out = z2 - 0.5 * (r2 - P.R_out * P.R_out) - P.R_char_sq * (P.ln_R_out - ln_r);
        return f_in >= 0.0 && f_out <= 0.0;
    }
    if (l.z < -kGeomEpsilon || l.z >= P.L + kGeomEpsilon) return false;
    const double r2 = l.x * l.x + l.y * l.y;
    const double Re = P.R + kGeomEpsilon;
    return r2 <= Re * Re;
}

// trilinear interpolation with clamping (fieldArrayLoader.cpp:103-162) on a
// uniform grid: the axis search becomes an index computation.
__device__ __forceinline__ void locate_f64(double v, double v0, double v_last, double inv_d,
                                           double d, int n, int& i0, double& a) {
    if (v <= v0) {
        i0 = 0;
        a = 0.0;
    } else if (v >= v_last) {
        i0 = n - 2;
        a = 1.0;
    } else {
        int i = static_cast<int>((v - v0) * inv_d);
        i = i > n - 2 ? n - 2 : (i < 0 ? 0 : i);
        i0 = i;
        a = (v - (v0 + static_cast<double>(i) * d)) * inv_d;
    }
}

// interpolate_field's axis search (fieldArrayLoader.cpp:115-137) on a
// non-uniform axis: clamp below / above, else the first interval i with
// ax[i] <= v <= ax[i+1] - for a non-decreasing axis that is the first i with
// ax[i+1] >= v, found by bisection; NaN takes the search's fallback.
__device__ __forceinline__ void locate_axis_f64(const double* __restrict__ ax, int n, double v, int& i0,
                                                double& a) {
    if (v <= ax[0]) {
        i0 = 0;
        a = 0.0;
    } else if (v >= ax[n - 1] || v != v) {
        i0 = n - 2;
        a = 1.0;
    } else {
        int first = 0, len = n - 1;
        while (len > 0) {
            const int half = len >> 1;
            if (ax[first + 1 + half] < v) {
                first += half + 1;
                len -= half + 1;
            } else {
                len = half;
            }
        }
        i0 = first;
        const double dv = ax[first + 1] - ax[first];
        a = (dv > 0.0) ? (v - ax[first]) / dv : 0.0;
    }
}

// Time factor of field-array term j (CompositeFieldProvider::compute_scale_factor,
// CompositeFieldProvider.h:117-147; the host resolves each kind to a, w, phi).
// Scale of field-array term j at a stage: t is the stage time, or with voltage
// waveforms the row of the host's scale table (wave-uniform either way).
__device__ __forceinline__ double term_scale_f64(const StepParams& P, int j, double t) {
    if (P.term_scales) return P.term_scales[static_cast<size_t>(t) * kMaxFieldTerms + j];
    return P.term_rf[j] ? P.term_a[j] * cos(P.term_w[j] * t + P.term_phi[j]) : P.term_a[j];
}

// GridFieldProvider (one term) or CompositeFieldProvider::get_E
// (CompositeFieldProvider.h:80-97: per term, interpolate, then add scale * E_i
// in term order); t is the stage time.  GEN = false: uniform grid, one term.
template <bool GEN>
__device__ __forceinline__ d3 field_array_f64(const StepParams& P, d3 p, double t) {
    int ix, iy, iz;
    double ax, ay, az;
    if (GEN && P.ax_x) {
        locate_axis_f64(P.ax_x, P.nx, p.x, ix, ax);
        locate_axis_f64(P.ax_y, P.ny, p.y, iy, ay);
        locate_axis_f64(P.ax_z, P.nz, p.z, iz, az);
    } else {
        locate_f64(p.x, P.x0, P.x_last, P.inv_dx, P.dx, P.nx, ix, ax);
        locate_f64(p.y, P.y0, P.y_last, P.inv_dy, P.dy, P.ny, iy, ay);
        locate_f64(p.z, P.z0, P.z_last, P.inv_dz, P.dz, P.nz, iz, az);
    }
    const size_t sx = 1, sy = static_cast<size_t>(P.nx), sz = static_cast<size_t>(P.nx) * P.ny;
    const size_t base = ix + sy * iy + sz * iz;
    const int nt = (GEN && P.n_terms > 1) ? P.n_terms : 1;
    d3 r{0.0, 0.0, 0.0};
    for (int j = 0; j < nt; ++j) {
        d3 e{0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int dx = c >> 2, dy = (c >> 1) & 1, dzz = c & 1;
            const double w = (dx ? ax : 1.0 - ax) * (dy ? ay : 1.0 - ay) * (dzz ? az : 1.0 - az);
            const size_t idx = (base + dx * sx + dy * sy + dzz * sz) * nt + j;
            double ex, ey, ez;
            if (P.fa_f64) {
                const double4 v = P.fa_f64[idx];
                ex = v.x; ey = v.y; ez = v.z;
            } else {
                const float4 v = P.fa_f32[idx];
                ex = v.x; ey = v.y; ez = v.z;
            }
            e.x += w * ex;
            e.y += w * ey;
            e.z += w * ez;
        }
        if (nt == 1) {
            r = e;
        } else {
            const double s = term_scale_f64(P, j, t);
            r.x += s * e.x;
            r.y += s * e.y;
            r.z += s * e.z;
        }
    }
    if (P.fa_scale != 1.0) r = r * P.fa_scale;
    return r;
}

// c_rf: the stage's RF coefficient (analytic fields) or stage time (field
// arrays); ez: the stage's axial field (analytic IMS / quadrupole fields).
template <int FIELD>
__device__ __forceinline__ d3 field_f64(const StepParams& P, d3 p, double c_rf, double ez) {
    if constexpr (FIELD == FIELD_LINEAR) {
        // AnalyticalFieldModel.cpp:136-161 (IMS) and :219-248 (quadrupole)
        d3 E{0.0, 0.0, ez};
        if (P.has_rf) {
            E.x = -2.0 * p.x * c_rf;
            E.y = 2.0 * p.y * c_rf;
        }
        return E;
    } else if constexpr (FIELD == FIELD_ORBITRAP) {
        // AnalyticalFieldModel.cpp:188-217
        d3 l{p.x - P.origin[0], p.y - P.origin[1], p.z - P.origin[2]};
        if (!P.rot_identity) l = mat_mul(P.g2l, l);
        const double r_sq = fmax(1e-18, l.x * l.x + l.y * l.y);
        const double C = 1.0 - P.R_char_sq / r_sq;
        const double hk = 0.5 * P.orbi_k;
        d3 E{hk * l.x * C, hk * l.y * C, -P.orbi_k * l.z};
        if (!P.rot_identity) E = mat_mul(P.l2g, E);
        return E;
    } else {
        // c_rf carries the stage time here
        return field_array_f64<FIELD == FIELD_ARRAY_X>(P, p, c_rf);
    }
}

// DampingForce::calculate_gamma (DampingForce.cpp:105-277) with the per-ion
// factor hoisted: Langevin gamma = lpre / max(1e-6, |v|); Friction constant.
__device__ __forceinline__ double gamma_f64(const StepParams& P, double lpre, d3 v) {
    if (P.damping_model == 3) return lpre / fmax(1e-6, sqrt(dot(v, v)));
    return lpre;
}

// EnvironmentConfig::gas_velocity_at (EnvironmentConfig.h:96-129) at the
// global position p: the constant gas velocity, or with the AxialParabolic
// flow l2g (0, 0, u_max max(0, 1 - r^2 / R^2)), r the radial distance of
// g2l (p - origin) (u_max everywhere when the domain has no radius).
__device__ __forceinline__ d3 gas_velocity_f64(const StepParams& P, d3 p) {
    if (!P.flow_parabolic) return d3{P.u_gas[0], P.u_gas[1], P.u_gas[2]};
    double uz = P.flow_umax;
    if (P.flow_R2 > 0.0) {
        const double dx = p.x - P.origin[0], dy = p.y - P.origin[1], dz = p.z - P.origin[2];
        const double lx = P.g2l[0] * dx + P.g2l[1] * dy + P.g2l[2] * dz;
        const double ly = P.g2l[3] * dx + P.g2l[4] * dy + P.g2l[5] * dz;
        uz = P.flow_umax * fmax(0.0, 1.0 - (lx * lx + ly * ly) / P.flow_R2);
    }
    return d3{P.l2g[2] * uz, P.l2g[5] * uz, P.l2g[8] * uz};
}

// FLOW: the gas velocity at p (gas_velocity_f64), else the constant one.
template <int FIELD, int COLL, bool FLOW = false>
__device__ __forceinline__ d3 force_f64(const StepParams& P, d3 p, d3 v, double c_rf, double ez,
                                        double q, double m, double lpre) {
    d3 F = field_f64<FIELD>(P, p, c_rf, ez) * q;
    if constexpr (COLL == COLL_DAMPING) {
        const double g = gamma_f64(P, lpre, v);
        if (g > 0.0) {
            const d3 u = FLOW ? gas_velocity_f64(P, p) : d3{P.u_gas[0], P.u_gas[1], P.u_gas[2]};
            const d3 rel{v.x - u.x, v.y - u.y, v.z - u.z};
            F = F + rel * (-g * m);
        }
    }
    return F;
}

__device__ __forceinline__ d3 iso_dir_f64(PhiloxStream& rng) {
    const double u1 = rng.uniform_f64();
    const double u2 = rng.uniform_f64();
    const double cosT = 2.0 * u1 - 1.0;
    const double sinT = sqrt(fmax(0.0, 1.0 - cosT * cosT));
    const double phi = 2.0 * kPi * u2;
    double s, c;
    sincos(phi, &s, &c);
    return d3{sinT * c, sinT * s, cosT};
}

__device__ __forceinline__ void ortho_basis_f64(d3 e, d3& t1, d3& t2) {
    const d3 a = (fabs(e.x) < 0.9) ? d3{1, 0, 0} : d3{0, 1, 0};
    d3 c1 = cross(e, a);
    double n1 = sqrt(dot(c1, c1));
    t1 = n1 > 0.0 ? d3{c1.x / n1, c1.y / n1, c1.z / n1} : d3{0, 0, 0};
    d3 c2 = cross(e, t1);
    double n2 = sqrt(dot(c2, c2));
    t2 = n2 > 0.0 ? d3{c2.x / n2, c2.y / n2, c2.z / n2} : d3{0, 0, 0};
}

// CollisionKernels::ehss_collision_with_orientation (CollisionKernels.cpp:44-137,
// 241-281, 283-381): the molecule in orientation O (row-major 3x3, in the
// collision-axis frame), the impact range widened to sqrt(sigma_eff / pi)
// when that is larger, first-contact ray test over up to max_attempts
// impact parameters, specular reflection.  Rotated atoms are recomputed per
// attempt instead of stored (same arithmetic).
template <class Rot>
__device__ d3 ehss_core_f64(const StepParams& P, d3 v_ion, d3 v_n, double mi, double mn, double r_gas, uint32_t a0,
                            uint32_t nat, const Rot& O, double sigma_eff, int max_attempts, PhiloxStream& rng) {
    d3 vrel = v_ion - v_n;
    double vrel_mag = sqrt(dot(vrel, vrel));
    if (vrel_mag <= kMinVelocityMag) {
        vrel = iso_dir_f64(rng) * kMinVelocityMag;
        vrel_mag = kMinVelocityMag;
    }
    const d3 ax{vrel.x / vrel_mag, vrel.y / vrel_mag, vrel.z / vrel_mag};
    d3 t1, t2;
    ortho_basis_f64(ax, t1, t2);
    const double Bm[3][3] = {{t1.x, t2.x, ax.x}, {t1.y, t2.y, ax.y}, {t1.z, t2.z, ax.z}};
    double Rm[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Rm[3 * i + j] = Bm[i][0] * O[j] + Bm[i][1] * O[3 + j] + Bm[i][2] * O[6 + j];
    auto atom = [&](uint32_t j) -> double4 {
        const float4 f = P.atoms[a0 + j];
        return make_double4(f.x, f.y, f.z, f.w);
    };
    double b_max = 0.0;
    for (uint32_t j = 0; j < nat; ++j) {
        const double4 c = atom(j);
        const d3 ra = mat_mul(Rm, d3{c.x, c.y, c.z});
        const double sra = dot(ax, ra);
        const d3 rp = ra - ax * sra;
        const double cur = sqrt(dot(rp, rp)) + c.w;
        if (cur > b_max) b_max = cur;
    }
    b_max += r_gas;
    if (sigma_eff > 0.0) {
        const double b_sigma = sqrt(sigma_eff / kPi);
        if (b_sigma > b_max) b_max = b_sigma;
    }
    for (int attempt = 0; attempt < max_attempts; ++attempt) {
        const double uA = rng.uniform_f64();
        const double uB = rng.uniform_f64();
        const double b = b_max * sqrt(uA);
        double sphi, cphi;
        sincos(2.0 * kPi * uB, &sphi, &cphi);
        const d3 off = t1 * (b * cphi) + t2 * (b * sphi);
        bool hit = false;
        d3 nb{0, 0, 0};
        double best_s = 0.0;
        for (uint32_t j = 0; j < nat; ++j) {
            const double4 c = atom(j);
            const d3 ra = mat_mul(Rm, d3{c.x, c.y, c.z});
            const double Rsum = c.w + r_gas;
            const double Rsum2 = Rsum * Rsum;
            const d3 rel = off - ra;
            const double sstar = dot(ax, rel);
            const d3 dv = rel - ax * sstar;
            const double dmin2 = dot(dv, dv);
            if (dmin2 <= Rsum2) {
                const double h = sqrt(fmax(0.0, Rsum2 - dmin2));
                const double s_hit = sstar - h;
                const d3 ph = rel - ax * s_hit;
                const double phr2 = dot(ph, ph);
                d3 nc;
                if (phr2 <= kMinContactDistSq) {
                    nc = ax * -1.0;
                } else {
                    nc = ph * (1.0 / sqrt(phr2));
                    if (dot(ax, nc) > 0.0) nc = nc * -1.0;
                }
                if (!hit || s_hit < best_s) {
                    hit = true;
                    nb = nc;
                    best_s = s_hit;
                }
            }
        }
        if (hit) {
            const double vdn = dot(vrel, nb);
            const d3 vr = vrel - nb * (2.0 * vdn);
            const double inv_mt = 1.0 / (mi + mn);
            const d3 Vcom = (v_ion * mi + v_n * mn) * inv_mt;
            return Vcom + vr * (mn * inv_mt);
        }
        if (attempt == kBmaxExpansionAttempt) b_max *= kBmaxExpansionFactor;
    }
    return v_ion;
}

// CollisionKernels::ehss_collision (CollisionKernels.cpp:13-42): a Shoemake
// orientation (CollisionGeometry::generate_random_rotation), then the core
// with sigma_eff = 0 and 256 attempts; r_gas = the contact radius.
__device__ d3 ehss_f64(const StepParams& P, d3 v_ion, d3 v_n, double mi, double mn, double r_gas,
                       uint32_t a0, uint32_t nat, PhiloxStream& rng) {
    const double u1 = rng.uniform_f64();
    const double u2 = rng.uniform_f64();
    const double u3 = rng.uniform_f64();
    if (nat == 0) return v_ion;
    const double q1 = sqrt(1.0 - u1) * sin(2.0 * kPi * u2);
    const double q2 = sqrt(1.0 - u1) * cos(2.0 * kPi * u2);
    const double q3 = sqrt(u1) * sin(2.0 * kPi * u3);
    const double q4 = sqrt(u1) * cos(2.0 * kPi * u3);
    const double x = q1, y = q2, z = q3, w = q4;
    double O[9];
    O[0] = 1 - 2 * (y * y + z * z); O[1] = 2 * (x * y - z * w);     O[2] = 2 * (x * z + y * w);
    O[3] = 2 * (x * y + z * w);     O[4] = 1 - 2 * (x * x + z * z); O[5] = 2 * (y * z - x * w);
    O[6] = 2 * (x * z - y * w);     O[7] = 2 * (y * z + x * w);     O[8] = 1 - 2 * (x * x + y * y);
    return ehss_core_f64(P, v_ion, v_n, mi, mn, r_gas, a0, nat, O, 0.0, kEhssMaxAttempts, rng);
}

// One collision stage; returns true on a collision event.  FLOW: the neutrals
// move with ug (the gas velocity at the ion's position), else with the
// constant gas velocity.
template <int COLL, bool FLOW = false>
__device__ __forceinline__ bool collide_f64(const StepParams& P, const SpeciesDev& sp, uint32_t sidx,
                                            double m, double lpre, d3& v, PhiloxStream& rng, double dt,
                                            d3 ug = d3{0.0, 0.0, 0.0}) {
    const d3 u_gas = FLOW ? ug : d3{P.u_gas[0], P.u_gas[1], P.u_gas[2]};
    if constexpr (COLL == COLL_EHSS_MIX) {
        // EHSSCollisionHandler::handle_collision mixture path
        // (EHSSCollisionHandler.cpp:311-424): every component visited in order,
        // a neutral drawn for each participating one, rates from the actual
        // relative speeds, one test on the total rate, the partner drawn with
        // weight k_c (the last component when the running sum falls short),
        // then the geometry method with the partner's neutral, mass and contact
        // radius and no second test.  Offline sets are not used with a mixture.
        const double* sg = P.emix_sigma + static_cast<size_t>(sidx) * P.n_emix;
        d3 vn[kMaxGasComponents];
        double k[kMaxGasComponents];
        double k_total = 0.0;
#pragma unroll
        for (int c = 0; c < kMaxGasComponents; ++c) {
            k[c] = 0.0;
            vn[c] = d3{0.0, 0.0, 0.0};
            if (c < P.n_emix && P.emix_part[c]) {
                const double sn = P.emix_sig[c];
                double u1 = rng.uniform_f64(), u2 = rng.uniform_f64();
                vn[c].x = sn * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
                u1 = rng.uniform_f64(); u2 = rng.uniform_f64();
                vn[c].y = sn * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
                u1 = rng.uniform_f64(); u2 = rng.uniform_f64();
                vn[c].z = sn * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
                vn[c] = vn[c] + u_gas;
                const d3 vr = v - vn[c];
                const double vm = sqrt(dot(vr, vr));
                if (sg[c] > 0.0 && P.emix_n[c] > 0.0 && vm >= kMinRelativeVelocity) {
                    k[c] = P.emix_n[c] * sg[c] * vm;
                    k_total += k[c];
                }
            }
        }
        if (!(k_total > 0.0)) return false;
        double Pt = 1.0;
        if (k_total * dt <= 50.0) Pt = 1.0 - exp(-k_total * dt);
        if (rng.uniform_f64() >= Pt) return false;
        const double r = rng.uniform_f64() * k_total;
        int sel = -1;
        double cum = 0.0;
#pragma unroll
        for (int c = 0; c < kMaxGasComponents; ++c) {
            if (c < P.n_emix) {
                cum += k[c];
                if (sel < 0 && r < cum) sel = c;
            }
        }
        if (sel < 0) sel = P.n_emix - 1;
        d3 vs{0.0, 0.0, 0.0};
        double ms = 0.0, rs = 0.0, sig_eff = 0.0;
#pragma unroll
        for (int c = 0; c < kMaxGasComponents; ++c) {
            if (c == sel) {
                vs = vn[c];
                ms = P.emix_m[c];
                rs = P.emix_r[c];
                sig_eff = sg[c];
            }
        }
        if (sig_eff <= 0.0) return false;
        const d3 vr = v - vs;
        if (sqrt(dot(vr, vr)) < kMinRelativeVelocity) return false;
        v = ehss_f64(P, v, vs, m, ms, rs, sp.atom_offset, sp.atom_count, rng);
        return true;
    }
    if constexpr (COLL == COLL_HSS_MIX) {
        {
            // HSSCollisionHandler::handle_collision mixture path
            // (HSSCollisionHandler.cpp:105-249): a neutral per component, rates
            // from the actual relative speeds, one test on the total rate, the
            // partner drawn with weight k_c, HSS with the neutral drawn for it
            const double* sg = P.mix_sigma + static_cast<size_t>(sidx) * P.n_mix;
            d3 vn[kMaxGasComponents];
            double k[kMaxGasComponents];
            double k_total = 0.0;
#pragma unroll
            for (int c = 0; c < kMaxGasComponents; ++c) {
                k[c] = -1.0;
                vn[c] = d3{0.0, 0.0, 0.0};
                if (c < P.n_mix) {
                    const double sn = P.mix_sig[c];
                    double u1 = rng.uniform_f64(), u2 = rng.uniform_f64();
                    vn[c].x = sn * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
                    u1 = rng.uniform_f64(); u2 = rng.uniform_f64();
                    vn[c].y = sn * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
                    u1 = rng.uniform_f64(); u2 = rng.uniform_f64();
                    vn[c].z = sn * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
                    vn[c] = vn[c] + u_gas;
                    const d3 vr = v - vn[c];
                    const double vm = sqrt(dot(vr, vr));
                    if (vm >= kMinRelativeVelocity) {
                        k[c] = P.mix_n[c] * sg[c] * vm;
                        k_total += k[c];
                    }
                }
            }
            if (!(k_total > 0.0)) return false;
            double Pt = 1.0;
            if (k_total * dt <= 50.0) Pt = 1.0 - exp(-k_total * dt);
            if (rng.uniform_f64() >= Pt) return false;
            const double r = rng.uniform_f64() * k_total;
            // the partner: the first component whose running sum exceeds r,
            // else the first one with a rate
            int sel = -1, first = -1;
            double cum = 0.0;
#pragma unroll
            for (int c = 0; c < kMaxGasComponents; ++c) {
                if (k[c] >= 0.0) {
                    if (first < 0) first = c;
                    cum += k[c];
                    if (sel < 0 && r < cum) sel = c;
                }
            }
            if (sel < 0) sel = first;
            d3 vs{0.0, 0.0, 0.0};
            double ms = 0.0;
#pragma unroll
            for (int c = 0; c < kMaxGasComponents; ++c) {
                if (c == sel) {
                    vs = vn[c];
                    ms = P.mix_m[c];
                }
            }
            // CollisionKernels::hss_collision (CollisionKernels.cpp:139-169)
            d3 vr = v - vs;
            double vm = sqrt(dot(vr, vr));
            if (vm <= kMinVelocityMag) {
                vr = iso_dir_f64(rng) * kMinVelocityMag;
                vm = kMinVelocityMag;
            }
            const d3 nd = iso_dir_f64(rng);
            const double inv_mt = 1.0 / (m + ms);
            const d3 Vcom = (v * m + vs * ms) * inv_mt;
            v = Vcom + (nd * vm) * (ms * inv_mt);
            return true;
        }
    }
    if constexpr (COLL == COLL_HSS || COLL == COLL_EHSS || COLL == COLL_EHSS_TAB) {
        double sigma = (COLL == COLL_HSS) ? sp.sigma_hss : sp.sigma_ehss;
        // EHSS offline sample set of the species (EHSSCollisionHandler.cpp:151-172):
        // the orientation is drawn first and its sigma_eff enters the test
        uint4 off = make_uint4(0u, 0u, 0u, 0u);
        uint2 ori = make_uint2(0u, 0u);
        uint32_t oi = 0;
        if constexpr (COLL == COLL_EHSS_TAB) {
            if (P.ehss_off) {
                off = P.ehss_off[sidx];
                if (off.y > 0u) {
                    oi = static_cast<uint32_t>(rng.uniform_f64() * static_cast<double>(off.y));
                    if (oi >= off.y) oi = off.y - 1u;
                    sigma = P.ehss_off_sigma[off.x + oi];
                }
            }
            // else the species' orientation samples (EHSSCollisionHandler.cpp:
            // 187-217): an orientation drawn first, its area for the gas as sigma_eff
            if (off.y == 0u && P.ehss_ori) {
                ori = P.ehss_ori[sidx];
                if (ori.y > 0u) {
                    oi = static_cast<uint32_t>(rng.uniform_f64() * static_cast<double>(ori.y));
                    if (oi >= ori.y) oi = ori.y - 1u;
                    sigma = P.ehss_ori_area[ori.x + oi];
                }
            }
        }
        if (sigma <= 0.0) return false;
        // VelocitySampling::sample_neutral_velocity (VelocitySampling.cpp:9-35)
        d3 vn;
        {
            double u1 = rng.uniform_f64(), u2 = rng.uniform_f64();
            vn.x = P.sigma_neutral * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
            u1 = rng.uniform_f64(); u2 = rng.uniform_f64();
            vn.y = P.sigma_neutral * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
            u1 = rng.uniform_f64(); u2 = rng.uniform_f64();
            vn.z = P.sigma_neutral * sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
            vn = vn + u_gas;
        }
        const d3 vrel = v - vn;
        const double vrel_mag = sqrt(dot(vrel, vrel));
        if (vrel_mag < kMinRelativeVelocity) return false;
        const double Pc = 1.0 - exp(-P.n * sigma * vrel_mag * dt);
        if (rng.uniform_f64() >= Pc) return false;
        if constexpr (COLL == COLL_HSS) {
            // CollisionKernels::hss_collision (CollisionKernels.cpp:139-169)
            d3 vr = vrel;
            double vm = vrel_mag;
            if (vm <= kMinVelocityMag) {
                vr = iso_dir_f64(rng) * kMinVelocityMag;
                vm = kMinVelocityMag;
            }
            const d3 nd = iso_dir_f64(rng);
            const double inv_mt = 1.0 / (m + P.m_gas);
            const d3 Vcom = (v * m + vn * P.m_gas) * inv_mt;
            v = Vcom + (nd * vm) * (P.m_gas * inv_mt);
        } else if (COLL == COLL_EHSS_TAB && off.y > 0u) {
            // offline branch of apply_collision (EHSSCollisionHandler.cpp:253-277)
            uint32_t mi = static_cast<uint32_t>(rng.uniform_f64() * static_cast<double>(off.z));
            if (mi >= off.z) mi = off.z - 1u;
            double mu = P.ehss_off_mu[static_cast<size_t>(off.w) + static_cast<size_t>(oi) * off.z + mi];
            mu = fmin(1.0, fmax(-1.0, mu));
            const double sin_theta = sqrt(fmax(0.0, 1.0 - mu * mu));
            const double phi = 2.0 * kPi * rng.uniform_f64();
            const d3 ax{vrel.x / vrel_mag, vrel.y / vrel_mag, vrel.z / vrel_mag};
            d3 t1, t2;
            ortho_basis_f64(ax, t1, t2);
            const d3 nc = t1 * (sin_theta * cos(phi)) + t2 * (sin_theta * sin(phi)) + ax * (-mu);
            const double vdn = vrel.x * nc.x + vrel.y * nc.y + vrel.z * nc.z;
            const d3 vr = vrel - nc * (2.0 * vdn);
            const double inv_mt = 1.0 / (m + P.m_gas);
            const d3 Vcom = (v * m + vn * P.m_gas) * inv_mt;
            v = Vcom + vr * (P.m_gas * inv_mt);
        } else if (COLL == COLL_EHSS_TAB && ori.y > 0u) {
            // the stored orientation, sigma_eff and the forced-hit budget
            // (EHSSCollisionHandler.cpp:286-293, CollisionKernels.cpp:110)
            const double* O = P.ehss_ori_R + 9ull * (ori.x + oi);
            v = ehss_core_f64(P, v, vn, m, P.m_gas, P.r_gas, sp.atom_offset, sp.atom_count, O, sigma,
                              kEhssMaxAttempts * 16, rng);
        } else {
            v = ehss_f64(P, v, vn, m, P.m_gas, P.r_gas, sp.atom_offset, sp.atom_count, rng);
        }
        return true;
    } else if constexpr (COLL == COLL_DAMPING) {
        if (P.ou_kicks) {
            // ou_velocity_update, apply_damping = false (CollisionKernels.cpp:204-213)
            const double g = gamma_f64(P, lpre, v);
            if (g <= 0.0 || dt <= 0.0) return false;
            const double kBT_over_m = kBoltzmann * P.T / m;
            const double s = sqrt(2.0 * g * kBT_over_m * dt);
            // three normals of the stream (a pair's second value carries over
            // to the next collision substep)
            v.x += s * rng.normal_f64();
            v.y += s * rng.normal_f64();
            v.z += s * rng.normal_f64();
        }
        return false;
    } else {
        return false;
    }
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void add_counters_to(unsigned long long* counters, int stripes, uint32_t pushed,
                                                uint32_t coll) {
    const uint32_t sp = wave_sum_u32(pushed);
    const uint32_t sc = wave_sum_u32(coll);
    if ((threadIdx.x & 63) == 0 && (sp | sc)) {
        uint32_t s = 0;
        if (stripes > 1)
            s = ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % static_cast<uint32_t>(stripes)) *
                kCounterStride;
        if (sp) atomicAdd(&counters[s], static_cast<unsigned long long>(sp));
        if (sc) atomicAdd(&counters[s + 1], static_cast<unsigned long long>(sc));
    }
}

__device__ __forceinline__ void add_counters(const StepParams& P, uint32_t pushed, uint32_t coll) {
    add_counters_to(P.counters, P.counter_stripes, pushed, coll);
}

}  // namespace icarion
