// step_params.hpp - parameter block of the fused ion-step kernels, filled on
// the host from the C-ABI descriptors and passed by value as the kernel
// argument (so every field is a wave-uniform scalar-register value).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/icarion_gpu.h"

namespace icarion {

// Field kinds the kernels are specialised for.
enum FieldKind : int {
    FIELD_LINEAR = 0,    // IMS / TIMS / quadrupole: (-2x c(t), 2y c(t), Ez)
    FIELD_ORBITRAP = 1,  // hyperlogarithmic Orbitrap field in local coords
    FIELD_ARRAY = 2,     // trilinear field array: uniform grid, one term
    FIELD_ARRAY_X = 3    // trilinear field array: non-uniform axes and/or
                         // several time-scaled terms
};
enum CollKind : int {
    COLL_NONE = 0,
    COLL_HSS = 1,
    COLL_EHSS = 2,
    COLL_DAMPING = 3,    // Langevin or Friction damping force (+ OU kicks)
    COLL_HSS_MIX = 4,    // HSS over a gas mixture
    COLL_EHSS_TAB = 5,   // EHSS with offline sample sets for some species
    COLL_EHSS_MIX = 6    // EHSS over a gas mixture
};

// Per-species record on the device.
struct SpeciesDev {
    double sigma_hss;
    double sigma_ehss;
    double K0_cm2_Vs;
    uint32_t atom_offset;
    uint32_t atom_count;
    float feat[3];           // surrogate cluster features
    float pad;
};

// Launch counters of a single cloud are striped over cache lines so that the
// waves' atomics (which on a multi-XCD device are served beyond the XCDs' L2)
// do not serialise on one address: stripe s holds {ion-steps, collisions} at
// counters[s * kCounterStride].
constexpr int kCounterStripes = 64;
constexpr int kCounterStride = 16;   // 128 B

// Field-array terms a domain may superpose (CompositeFieldProvider).
constexpr int kMaxFieldTerms = 4;
// Components of an HSS gas mixture.
constexpr int kMaxGasComponents = 4;

// Per-step uniform values at the step start, mid point and end: for analytic
// fields the RF coefficient c = U_eff(t)/r0^2 and the axial field E_z(t); for a
// field array of several time-scaled terms c is the stage time itself (the
// kernel evaluates the scales).  nrel: the gas number density of the step over
// the domain's constant one (1 unless the pressure follows a waveform); every
// density-proportional rate of the step (collision rates of each gas
// component, the damping coefficient) is scaled by it.
struct StepEntry {
    double t;
    double c0;
    double ch;
    double c1;
    double e0;
    double eh;
    double e1;
    double nrel;
};

struct IonPtrs {
    double *x, *y, *z, *vx, *vy, *vz;
    const double *mass, *charge;
    const uint32_t *species;
    uint8_t *flags;          // bit 0 active, bit 1 born
    const double *birth;     // may be null: born at the first step
    double *death;
    double *arrival;
    uint32_t *ncoll;
    int32_t *domain;         // several domains: each ion's domain index
};

// RK45 adaptive step control (RK45Strategy::AdaptiveConfig, RK45Strategy.h:54-65).
struct Rk45Cfg {
    double atol, rtol, safety, min_factor, max_factor, max_inc, max_dec, abs_min;
    int accept_at_dt_min, pad;
};

// Per-ion RK45 state resident on the device (RK45Strategy::RK45State,
// RK45Strategy.h:192-198, and the engine's dt_per_ion_) plus the engine record
// that carries the shared engine time from one step launch to the next.
struct Rk45Engine {
    double T;                       // engine time of the step about to run
    double dt_fallback;             // config dt (SimulationEngineUtils.cpp:82)
    unsigned long long step;        // absolute step number (RNG counter)
    unsigned long long tmax;        // ordered bits: max ion time after the step
    unsigned long long bmin;        // ordered bits: min birth > T of unborn ions
    unsigned long long dtmax;       // ordered bits: max accepted dt
    unsigned long long rejected;    // rejected trial steps
    unsigned long long failed;      // ions with no accepted trial
    unsigned long long bad_pressure;    // 1: the pressure waveform was <= 0 at T (T then stays)
};

struct Rk45Ptrs {
    double *t, *dt, *last_err, *k1x, *k1y, *k1z;
    uint8_t *fsal;
    Rk45Engine *eng;
};

// FP32 copies of the run constants used by the fast path, rounded once on the
// host (kernel arguments: wave-uniform, so they stay in scalar registers).
struct FastConst {
    float Ez, orbi_hk, orbi_k, R_char_sq;
    float ox, oy, oz;
    float u_gas[3];
    float flow_umax, flow_R2;   // AxialParabolic flow (StepParams.flow_*)
    float sigma_neutral;
    float n_dt;          // n * dt
    float m_gas;
    float L, half_L, R_eps_sq, R_in_sq, R_out, R_out_sq, ln_R_in, ln_R_out;
    float g2l[9], l2g[9];
};

struct StepParams {
    // geometry (DomainConfig.geometry)
    int instrument;
    int rot_identity;        // global->local rotation is the identity (origin shift still applies)
    int magnetic;
    int damping_model;       // 3 Langevin, 4 Friction (collision model code)
    int ou_kicks;
    int detect;              // arrival detection on
    int has_rf;
    int counter_stripes;     // 0: counters = {ion-steps, collisions}; k: k stripes of kCounterStride
    double origin[3];
    double L;
    double R;
    double R_in, R_out, R_char;
    double ln_R_in, ln_R_out;
    double g2l[9];
    double l2g[9];
    // fields
    double Ez_const;         // IMS: axial_V/L; quadrupole: axial_V/L if set
    double orbi_k;           // Orbitrap k
    double R_char_sq;
    double fa_scale;
    double B[3];
    double rf_U, rf_V, rf_omega, r0_sq;   // U_eff(t) = U + V cos(omega t) (RK45 evaluates it)
    // environment
    double T;
    double n;
    double m_gas;
    double r_gas;
    double alpha_gas;
    double u_gas[3];
    // EnvironmentConfig's AxialParabolic flow (flow_parabolic = 1): the gas
    // velocity at p is l2g (0, 0, flow_umax max(0, 1 - r^2 / flow_R2)), r the
    // local radial distance; flow_R2 = 0: flow_umax everywhere.  Read by the
    // extended (SUB) instantiations only.
    int flow_parabolic;
    double flow_umax, flow_R2;
    double sigma_neutral;    // sqrt(kB T / m_gas)
    double v_th;             // sqrt(8 kB T / (pi m_gas))
    double det_z;
    // field array (rectilinear node-centred grid, x fastest; with several
    // terms the node's terms are adjacent: element (node * n_terms + term))
    const float4 *fa_f32;
    const double4 *fa_f64;
    int nx, ny, nz, n_terms;
    const double *ax_x, *ax_y, *ax_z;     // non-uniform axes (NULL: uniform grid)
    const float *axf_x, *axf_y, *axf_z;   // the same axes in FP32
    // per term: scale = a (constant kinds) or a cos(w t + phi) (RF kind)
    int term_rf[kMaxFieldTerms];
    double term_a[kMaxFieldTerms], term_w[kMaxFieldTerms], term_phi[kMaxFieldTerms];
    // with voltage waveforms: every term's scale at every stage time, row r =
    // 3 * step + stage (0 start, 1 mid point, 2 end), kMaxFieldTerms per row;
    // the step table's stage values are then row numbers (NULL: closed form)
    const double *term_scales;
    // HSS gas mixture (n_mix = 0: the single gas above): the participating
    // components with n > 0, in order; sigma of species s with component c at
    // mix_sigma[s * n_mix + c]
    int n_mix;
    // collision substeps per engine step (>= 1): the collision stage runs
    // n_sub times with dt / n_sub on the pushed state, drawing from one
    // continuing stream (handle_collisions_cpu's subcycle loop,
    // SimulationEngineCollisionRuntime.cpp:84-191)
    int n_sub;
    double mix_n[kMaxGasComponents], mix_m[kMaxGasComponents];
    double mix_sig[kMaxGasComponents];   // sqrt(kB T / m_c)
    const double *mix_sigma;
    // EHSS gas mixture: every component in order (the EHSS handler's loop
    // visits them all), its participation flag, contact radius, and the EHSS
    // cross sections of species s at emix_sigma[s * n_emix + c]
    int n_emix;
    int field_array;    // several domains: this domain's E comes from its uniform field array
    int emix_part[kMaxGasComponents];
    double emix_n[kMaxGasComponents], emix_m[kMaxGasComponents];
    double emix_sig[kMaxGasComponents], emix_r[kMaxGasComponents];
    const double *emix_sigma;
    // offline EHSS sample sets per species (NULL: none): {first orientation,
    // orientations (0: the online geometry), mu per orientation, first mu}
    const uint4 *ehss_off;
    const double *ehss_off_sigma;   // sigma_eff per orientation
    const double *ehss_off_mu;      // mu samples, orientation-major
    // EHSS orientation samples per species (NULL: none): {first orientation,
    // orientations}; each orientation's rotation (row-major 3x3) and area
    const uint2 *ehss_ori;
    const double *ehss_ori_R;
    const double *ehss_ori_area;
    double x0, y0, z0;
    double dx, dy, dz;
    double inv_dx, inv_dy, inv_dz;
    double x_last, y_last, z_last;
    // species / atoms
    const SpeciesDev *species;
    const float4 *atoms;     // x, y, z, r of every EHSS atom, FP32
    const uint4 *mlp_blob;   // packed bf16 surrogate weights (surrogate.hpp)
    // per-step table, entries [0, n_steps]
    const StepEntry *table;
    // run
    double dt;
    uint64_t step0;
    uint64_t seed;
    uint32_t n_steps;
    uint32_t n_ions;
    // lane t of the launch takes ion order[t] (a permutation of the ions that
    // puts neighbours in the field-array grid on neighbouring lanes); null:
    // lane t takes ion t.  The ion keeps its index (state, random streams).
    const uint32_t *order;
    unsigned long long *counters;   // [0] ion-steps, [1] collisions
    FastConst fc;
    Rk45Cfg rk45;
    // RK45 with voltage waveforms on the analytic IMS / TIMS and quadrupole
    // fields: the slot descriptors (ARBITRARY points in device memory), the
    // mask of slots that hold one, and every slot's constant
    const icarion_waveform_desc *waves;
    uint32_t wave_mask;
    // RK45 with a pressure waveform (descriptor waves[ICARION_N_VOLTAGE_SLOTS]):
    // the engine step's density ratio is P(T) * pressure_scale
    uint32_t pressure_wave;
    double pressure_scale;
    double wave_const[ICARION_N_VOLTAGE_SLOTS];
};

// Fill P.fc from the FP64 fields (after P.dt is set).
inline void fill_fast_const(StepParams& P) {
    FastConst& C = P.fc;
    C.Ez = (float)P.Ez_const;
    C.orbi_k = (float)P.orbi_k;
    C.orbi_hk = (float)(0.5 * P.orbi_k);
    C.R_char_sq = (float)P.R_char_sq;
    C.ox = (float)P.origin[0]; C.oy = (float)P.origin[1]; C.oz = (float)P.origin[2];
    C.u_gas[0] = (float)P.u_gas[0]; C.u_gas[1] = (float)P.u_gas[1]; C.u_gas[2] = (float)P.u_gas[2];
    C.flow_umax = (float)P.flow_umax;
    C.flow_R2 = (float)P.flow_R2;
    C.sigma_neutral = (float)P.sigma_neutral;
    C.n_dt = (float)(P.n * P.dt);
    C.m_gas = (float)P.m_gas;
    C.L = (float)P.L;
    C.half_L = (float)(0.5 * P.L);
    C.R_eps_sq = (float)(P.R * P.R);
    C.R_in_sq = (float)(P.R_in * P.R_in);
    C.R_out = (float)P.R_out;
    C.R_out_sq = (float)(P.R_out * P.R_out);
    C.ln_R_in = (float)P.ln_R_in;
    C.ln_R_out = (float)P.ln_R_out;
    for (int c = 0; c < 9; ++c) {
        C.g2l[c] = (float)P.g2l[c];
        C.l2g[c] = (float)P.l2g[c];
    }
}

// ---- several domains ---------------------------------------------------------
constexpr int kMaxDomains = 8;

// One cylindrical domain's entry in DomainManager's z-sorted span list
// (DomainManager.cpp:43-63): z range and radius widened by 1e-12 m.
struct DomainSpan {
    double z_min, z_max, radius;
    int idx, pad;
};

struct MultiParams {
    const StepParams *dom;          // per domain; table = its coefficients, k0-based
    const DomainSpan *spans;        // sorted by z_min (n_spans; none with an Orbitrap)
    const StepEntry *t_table;       // engine times (any domain's table)
    const double *ap_sq;            // per domain: end aperture^2
    const SpeciesDev *species;
    unsigned long long *counters;
    int n_dom, n_spans, has_orbitrap, counter_stripes;
    double handoff_tol;             // shared_boundary_handoff_domain's tolerance
    double dt;
    uint64_t step0, seed;
    uint32_t k0, n_steps, n_ions;
    int n_sub;                      // collision substeps per step (every domain's n_sub)
    int has_field_array;            // some domain takes E from a field array
    int has_flow;                   // some domain has the AxialParabolic gas flow
};

// ---- batched clouds ----------------------------------------------------------
// One tile = up to blockDim ions of one cloud, advanced through all of that
// cloud's steps by one workgroup of the persistent batch kernel.
struct BatchTile {
    uint32_t cloud;
    uint32_t first;      // first ion of the tile within its cloud
};

// kind of a cloud's kernel body: integrator + 2 * field (0 linear, 1 Orbitrap)
// + 4 * collision (0 none, 1 HSS, 2 damping)
constexpr uint32_t kBatchKinds = 12;
// Precision classes of the batch, one persistent kernel each:
// 0 FP64, 1 FP32 + Kahan, 2 FP32 + Kahan with the EHSS surrogate.
constexpr int kBatchClasses = 3;

struct BatchParams {
    const StepParams *params;       // per cloud (n_steps = the cloud's steps)
    const IonPtrs *ions;            // per cloud
    const uint32_t *kinds;          // per cloud
    const BatchTile *tiles;         // in the order the queue hands them out
    uint32_t n_tiles;
    uint32_t with_mix;              // some cloud of the class has an HSS gas mixture
    uint32_t *queue;                // next tile to hand out
    unsigned long long *done;       // per cloud: clock at the end of its last tile
    unsigned long long *start;      // clock at the first workgroup's start
};

}  // namespace icarion
