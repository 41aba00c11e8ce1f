// step_device_f32.hpp - device functions of the FP32 + Kahan path shared by
// the fixed-step, surrogate, batch and RK45 kernels: compensated accumulation,
// hardware transcendentals, fields, containment, collisions and acceleration
// in FP32 (time and RF phase stay FP64).
#pragma once

#include "step_device.hpp"

namespace icarion {

using f3 = vec3<float>;

// compensated accumulation: (hi, lo) <- hi + lo + d  (Kahan / Fast2Sum)
__device__ __forceinline__ void kahan_add(float& hi, float& lo, float d) {
    const float y = d + lo;
    const float t = hi + y;
    lo = y - (t - hi);
    hi = t;
}

// hardware transcendentals: v_log_f32 (log2), v_exp_f32 (exp2), v_sin/cos_f32
// (argument in revolutions), v_sqrt_f32, v_rcp_f32
__device__ __forceinline__ float fast_ln(float u) { return __builtin_amdgcn_logf(u) * 0.69314718056f; }
__device__ __forceinline__ float fast_cos_rev(float r) { return __builtin_amdgcn_cosf(r); }
__device__ __forceinline__ float fast_sin_rev(float r) { return __builtin_amdgcn_sinf(r); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }


// c: the stage's RF coefficient (analytic fields) or the stage time (field
// arrays), in FP64 as the step table holds it.
template <int FIELD>
__device__ __forceinline__ f3 field_f32(const StepParams& P, const FastConst& C, f3 p, double c,
                                        double ez) {
    if constexpr (FIELD == FIELD_LINEAR) {
        const float c_rf = (float)c;
        f3 E{0.0f, 0.0f, (float)ez};
        if (P.has_rf) {
            E.x = -2.0f * p.x * c_rf;
            E.y = 2.0f * p.y * c_rf;
        }
        return E;
    } else if constexpr (FIELD == FIELD_ORBITRAP) {
        const f3 s{p.x - C.ox, p.y - C.oy, p.z - C.oz};
        f3 l = s;
        if (!P.rot_identity) {
            l = f3{C.g2l[0] * s.x + C.g2l[1] * s.y + C.g2l[2] * s.z,
                   C.g2l[3] * s.x + C.g2l[4] * s.y + C.g2l[5] * s.z,
                   C.g2l[6] * s.x + C.g2l[7] * s.y + C.g2l[8] * s.z};
        }
        const float r_sq = fmaxf(1e-18f, l.x * l.x + l.y * l.y);
        const float Cf = 1.0f - C.R_char_sq * __builtin_amdgcn_rcpf(r_sq);
        f3 E{C.orbi_hk * l.x * Cf, C.orbi_hk * l.y * Cf, -C.orbi_k * l.z};
        if (!P.rot_identity) {
            E = f3{C.l2g[0] * E.x + C.l2g[1] * E.y + C.l2g[2] * E.z,
                   C.l2g[3] * E.x + C.l2g[4] * E.y + C.l2g[5] * E.z,
                   C.l2g[6] * E.x + C.l2g[7] * E.y + C.l2g[8] * E.z};
        }
        return E;
    } else {
        // trilinear on float4 nodes; the axis search of the FP64 path on
        // non-uniform axes
        auto loc = [](float v, float v0, float vl, float inv_d, int n, int& i0, float& a) {
            if (v <= v0) { i0 = 0; a = 0.0f; }
            else if (v >= vl) { i0 = n - 2; a = 1.0f; }
            else {
                const float f = (v - v0) * inv_d;
                int i = static_cast<int>(f);
                i = i > n - 2 ? n - 2 : (i < 0 ? 0 : i);
                i0 = i;
                a = f - static_cast<float>(i);
            }
        };
        auto loc_axis = [](const float* __restrict__ ax, int n, float v, int& i0, float& a) {
            if (v <= ax[0]) { i0 = 0; a = 0.0f; }
            else if (v >= ax[n - 1] || v != v) { i0 = n - 2; a = 1.0f; }
            else {
                int first = 0, len = n - 1;
                while (len > 0) {
                    const int half = len >> 1;
                    if (ax[first + 1 + half] < v) { first += half + 1; len -= half + 1; }
                    else len = half;
                }
                i0 = first;
                const float dv = ax[first + 1] - ax[first];
                a = dv > 0.0f ? (v - ax[first]) / dv : 0.0f;
            }
        };
        constexpr bool GEN = FIELD == FIELD_ARRAY_X;
        int ix, iy, iz;
        float ax, ay, az;
        if (GEN && P.axf_x) {
            loc_axis(P.axf_x, P.nx, p.x, ix, ax);
            loc_axis(P.axf_y, P.ny, p.y, iy, ay);
            loc_axis(P.axf_z, P.nz, p.z, iz, az);
        } else {
            loc(p.x, (float)P.x0, (float)P.x_last, (float)P.inv_dx, P.nx, ix, ax);
            loc(p.y, (float)P.y0, (float)P.y_last, (float)P.inv_dy, P.ny, iy, ay);
            loc(p.z, (float)P.z0, (float)P.z_last, (float)P.inv_dz, P.nz, iz, az);
        }
        const size_t sy = static_cast<size_t>(P.nx), sz = static_cast<size_t>(P.nx) * P.ny;
        const size_t base = ix + sy * iy + sz * iz;
        f3 r{0.0f, 0.0f, 0.0f};
        if (!GEN || P.n_terms <= 1) {
#pragma unroll
            for (int c8 = 0; c8 < 8; ++c8) {
                const int dx = c8 >> 2, dy = (c8 >> 1) & 1, dzz = c8 & 1;
                const float w = (dx ? ax : 1.0f - ax) * (dy ? ay : 1.0f - ay) * (dzz ? az : 1.0f - az);
                const float4 e = P.fa_f32[base + dx + dy * sy + dzz * sz];
                r.x += w * e.x;
                r.y += w * e.y;
                r.z += w * e.z;
            }
        } else {
            // the time factors in FP64 (phase never in FP32), applied in FP32
            const int nt = P.n_terms;
            for (int j = 0; j < nt; ++j) {
                const float sj = (float)term_scale_f64(P, j, c);
                f3 e{0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int c8 = 0; c8 < 8; ++c8) {
                    const int dx = c8 >> 2, dy = (c8 >> 1) & 1, dzz = c8 & 1;
                    const float w = (dx ? ax : 1.0f - ax) * (dy ? ay : 1.0f - ay) * (dzz ? az : 1.0f - az);
                    const float4 v4 = P.fa_f32[(base + dx + dy * sy + dzz * sz) * nt + j];
                    e.x += w * v4.x;
                    e.y += w * v4.y;
                    e.z += w * v4.z;
                }
                r = r + e * sj;
            }
        }
        const float s = (float)P.fa_scale;
        if (s != 1.0f) r = r * s;
        return r;
    }
}

// Domain containment in FP32 (same predicates as contains_f64; the FP32
// position is good to ~1e-7 relative, far finer than any domain feature).
__device__ __forceinline__ bool contains_f32(const StepParams& P, const FastConst& C, float px,
                                             float py, float pz) {
    const f3 s{px - C.ox, py - C.oy, pz - C.oz};
    f3 l = s;
    if (!P.rot_identity) {
        l = f3{C.g2l[0] * s.x + C.g2l[1] * s.y + C.g2l[2] * s.z,
               C.g2l[3] * s.x + C.g2l[4] * s.y + C.g2l[5] * s.z,
               C.g2l[6] * s.x + C.g2l[7] * s.y + C.g2l[8] * s.z};
    }
    const float r2 = l.x * l.x + l.y * l.y;
    if (P.instrument == 3) {
        const float z_abs = fabsf(l.z);
        if (z_abs > C.half_L) return false;
        if (r2 > C.R_out_sq) return false;
        const float ln_r = 0.5f * fast_ln(r2);
        const float z2 = z_abs * z_abs;
        const float f_in = z2 - 0.5f * (r2 - C.R_in_sq) - C.R_char_sq * (C.ln_R_in - ln_r);
        const float f_out = z2 - 0.5f * (r2 - C.R_out * C.R_out) - C.R_char_sq * (C.ln_R_out - ln_r);
        return f_in >= 0.0f && f_out <= 0.0f;
    }
    if (l.z < 0.0f || l.z >= C.L) return false;
    return r2 <= C.R_eps_sq;
}

__device__ __forceinline__ f3 iso_dir_f32(PhiloxStream& rng) {
    const float u1 = rng.uniform_f32();
    const float u2 = rng.uniform_f32();
    const float cosT = 2.0f * u1 - 1.0f;
    const float sinT = fast_sqrt(fmaxf(0.0f, 1.0f - cosT * cosT));
    return f3{sinT * fast_cos_rev(u2), sinT * fast_sin_rev(u2), cosT};
}

// FLOW: the neutrals move with ug (the gas velocity at the ion's position),
// else with the constant gas velocity.
template <int COLL, bool FLOW = false>
__device__ __forceinline__ bool collide_f32(const StepParams& P, const FastConst& C,
                                            const SpeciesDev& sp, uint32_t sidx, float m,
                                            float inv_m_plus, float lpre, float kbt2_dt_over_m, f3& v,
                                            PhiloxStream& rng, float dt, float n_dt,
                                            f3 ug = f3{0.0f, 0.0f, 0.0f}) {
    const f3 u_gas = FLOW ? ug : f3{C.u_gas[0], C.u_gas[1], C.u_gas[2]};
    if constexpr (COLL == COLL_HSS_MIX) {
        {
            // the mixture path of collide_f64 in FP32 (rates, test and partner
            // choice as there; neutrals from two Box-Muller pairs)
            const double* sg = P.mix_sigma + static_cast<size_t>(sidx) * P.n_mix;
            f3 vn[kMaxGasComponents];
            float k[kMaxGasComponents];
            float k_total = 0.0f;
#pragma unroll
            for (int c = 0; c < kMaxGasComponents; ++c) {
                k[c] = -1.0f;
                vn[c] = f3{0.0f, 0.0f, 0.0f};
                if (c < P.n_mix) {
                    const float sn = (float)P.mix_sig[c];
                    const float u1 = rng.uniform_f32(), u2 = rng.uniform_f32();
                    const float u3 = rng.uniform_f32(), u4 = rng.uniform_f32();
                    const float r1 = sn * fast_sqrt(-2.0f * fast_ln(u1));
                    const float r2 = sn * fast_sqrt(-2.0f * fast_ln(u3));
                    vn[c] = f3{r1 * fast_cos_rev(u2) + u_gas.x, r1 * fast_sin_rev(u2) + u_gas.y,
                               r2 * fast_cos_rev(u4) + u_gas.z};
                    const f3 vr = v - vn[c];
                    const float vm = fast_sqrt(dot(vr, vr));
                    if (vm >= (float)kMinRelativeVelocity) {
                        k[c] = (float)(P.mix_n[c] * sg[c]) * vm;
                        k_total += k[c];
                    }
                }
            }
            if (!(k_total > 0.0f)) return false;
            const float x = k_total * dt;
            const float Pc = x < 1e-3f ? x * (1.0f - 0.5f * x * (1.0f - x * (1.0f / 3.0f)))
                                       : (x <= 50.0f ? 1.0f - __builtin_amdgcn_exp2f(-x * 1.44269504089f) : 1.0f);
            if (rng.uniform_f32() >= Pc) return false;
            const float r = rng.uniform_f32() * k_total;
            int sel = -1, first = -1;
            float cum = 0.0f;
#pragma unroll
            for (int c = 0; c < kMaxGasComponents; ++c) {
                if (k[c] >= 0.0f) {
                    if (first < 0) first = c;
                    cum += k[c];
                    if (sel < 0 && r < cum) sel = c;
                }
            }
            if (sel < 0) sel = first;
            f3 vs{0.0f, 0.0f, 0.0f};
            float ms = 0.0f;
#pragma unroll
            for (int c = 0; c < kMaxGasComponents; ++c) {
                if (c == sel) {
                    vs = vn[c];
                    ms = (float)P.mix_m[c];
                }
            }
            const f3 vr = v - vs;
            const float vm = fast_sqrt(dot(vr, vr));
            const f3 nd = iso_dir_f32(rng);
            const float inv_mt = 1.0f / (m + ms);
            const f3 Vcom = (v * m + vs * ms) * inv_mt;
            v = Vcom + nd * (vm * ms * inv_mt);
            return true;
        }
    } else if constexpr (COLL == COLL_HSS) {
        const float sigma = (float)sp.sigma_hss;
        if (!(sigma > 0.0f)) return false;
        // three Maxwellian components from two Box-Muller pairs (both branches)
        const float u1 = rng.uniform_f32(), u2 = rng.uniform_f32();
        const float u3 = rng.uniform_f32(), u4 = rng.uniform_f32();
        const float r1 = C.sigma_neutral * fast_sqrt(-2.0f * fast_ln(u1));
        const float r2 = C.sigma_neutral * fast_sqrt(-2.0f * fast_ln(u3));
        const f3 vn{r1 * fast_cos_rev(u2) + u_gas.x, r1 * fast_sin_rev(u2) + u_gas.y,
                    r2 * fast_cos_rev(u4) + u_gas.z};
        const f3 vrel = v - vn;
        const float vm2 = dot(vrel, vrel);
        const float vm = fast_sqrt(vm2);
        if (vm < (float)kMinRelativeVelocity) return false;
        // P = 1 - exp(-x), evaluated as -expm1(-x) to keep small x accurate
        const float x = n_dt * sigma * vm;
        const float Pc = x < 1e-3f ? x * (1.0f - 0.5f * x * (1.0f - x * (1.0f / 3.0f)))
                                   : 1.0f - __builtin_amdgcn_exp2f(-x * 1.44269504089f);
        if (rng.uniform_f32() >= Pc) return false;
        const f3 nd = iso_dir_f32(rng);
        const f3 Vcom = (v * m + vn * C.m_gas) * inv_m_plus;
        v = Vcom + nd * (vm * C.m_gas * inv_m_plus);
        return true;
    } else if constexpr (COLL == COLL_DAMPING) {
        if (P.ou_kicks) {
            const float g = (P.damping_model == 3) ? lpre / fmaxf(1e-6f, fast_sqrt(dot(v, v))) : lpre;
            if (!(g > 0.0f)) return false;
            const float s = fast_sqrt(g * kbt2_dt_over_m);
            const float u1 = rng.uniform_f32(), u2 = rng.uniform_f32();
            const float u3 = rng.uniform_f32(), u4 = rng.uniform_f32();
            const float r1 = s * fast_sqrt(-2.0f * fast_ln(u1));
            const float r2 = s * fast_sqrt(-2.0f * fast_ln(u3));
            v.x += r1 * fast_cos_rev(u2);
            v.y += r1 * fast_sin_rev(u2);
            v.z += r2 * fast_cos_rev(u4);
        }
        return false;
    } else {
        return false;
    }
}

// gas_velocity_f64 in FP32: the constant gas velocity, or the AxialParabolic
// profile at p.
__device__ __forceinline__ f3 gas_velocity_f32(const StepParams& P, const FastConst& C, f3 p) {
    if (!P.flow_parabolic) return f3{C.u_gas[0], C.u_gas[1], C.u_gas[2]};
    float uz = C.flow_umax;
    if (C.flow_R2 > 0.0f) {
        const float dx = p.x - C.ox, dy = p.y - C.oy, dz = p.z - C.oz;
        const float lx = C.g2l[0] * dx + C.g2l[1] * dy + C.g2l[2] * dz;
        const float ly = C.g2l[3] * dx + C.g2l[4] * dy + C.g2l[5] * dz;
        uz = C.flow_umax * fmaxf(0.0f, 1.0f - (lx * lx + ly * ly) / C.flow_R2);
    }
    return f3{C.l2g[2] * uz, C.l2g[5] * uz, C.l2g[8] * uz};
}

// FLOW: the gas velocity at p (gas_velocity_f32), else the constant one.
template <int FIELD, int COLL, bool FLOW = false>
__device__ __forceinline__ f3 accel_f32(const StepParams& P, const FastConst& C, f3 p, f3 v,
                                        double c, double ez, float qm, float lpre) {
    f3 a = field_f32<FIELD>(P, C, p, c, ez) * qm;
    if constexpr (COLL == COLL_DAMPING) {
        const float g = (P.damping_model == 3) ? lpre / fmaxf(1e-6f, fast_sqrt(dot(v, v))) : lpre;
        if (g > 0.0f) {
            const f3 u = FLOW ? gas_velocity_f32(P, C, p) : f3{C.u_gas[0], C.u_gas[1], C.u_gas[2]};
            a = a - f3{v.x - u.x, v.y - u.y, v.z - u.z} * g;
        }
    }
    return a;
}

}  // namespace icarion
