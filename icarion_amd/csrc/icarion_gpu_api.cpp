// icarion_gpu_api.cpp - C ABI (include/icarion_gpu.h) and host runtime of the
// gfx950 batch ion-step library.
//
// The handle keeps the ion ensemble resident in HBM as SoA arrays and drives
// the fused step kernels (step_kernels.hip) on one HIP stream.  Host-side work
// per icarion_gpu_step(): build the per-step table (engine time accumulated as
// t += dt exactly like the reference's t_i += dt, SimulationEngineLifecycle.cpp:172,
// and the RF coefficient U_eff(t)/r0^2 of AnalyticalFieldModel.cpp:226-241),
// upload it, and launch ceil(n_steps / steps_per_launch) kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/icarion_gpu.h"
#include "step_params.hpp"

#include <numeric>
#include <memory>
#include <random>

#include "surrogate.hpp"
#include "waveform.hpp"

#include <mutex>

// Scratch memory of icarion_gpu_step_noalloc, one copy per device held until
// the process ends: the striped launch counters, then the step table of up to
// ICARION_NOALLOC_MAX_STEPS entries.  In host code `scratch` names a host-side
// stand-in, not device memory: its device address is looked up per device
// (hipGetSymbolAddress, at handle creation with the device current) and only
// that address is handed to kernels and copies.
namespace {
constexpr size_t kNoAllocCounterBytes = size_t(icarion::kCounterStripes) * icarion::kCounterStride * 8;
constexpr size_t kNoAllocTableBytes = size_t(ICARION_NOALLOC_MAX_STEPS) * sizeof(icarion::StepEntry);
constexpr size_t kScratchBytes = kNoAllocCounterBytes + kNoAllocTableBytes;
static_assert(kNoAllocCounterBytes % 256 == 0, "the step table must start 256-byte aligned");
constexpr int kMaxScratchDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch[kMaxScratchDevices] = {};
std::string g_scratch_error[kMaxScratchDevices];
}  // namespace

namespace icarion {
hipError_t launch_step(int integrator, int field, int coll, int precision, int surrogate,
                       const StepParams& P, const IonPtrs& S, uint32_t block, hipStream_t stream);
hipError_t launch_rk45_step(int field, int coll, int precision, const StepParams& P, const IonPtrs& S,
                            const Rk45Ptrs& R, uint32_t block, hipStream_t stream);
hipError_t launch_rk45_init(const Rk45Ptrs& R, uint32_t n, double t0, double dt, hipStream_t stream);
hipError_t launch_pack_snapshot(const IonPtrs& S, uint32_t n, int full, double* pos, double* vel, uint32_t* spec,
                                int32_t* dom, uint8_t* flags, uint8_t* prev, void* rec, uint32_t* count,
                                uint32_t* host_count, uint32_t* next_count, hipStream_t s);
hipError_t launch_zero_u32(uint32_t* p, hipStream_t s);
uint32_t arrival_blocks(uint32_t n);
hipError_t launch_arrival_partials(const double* arrival, uint32_t n, void* partials, hipStream_t s);
hipError_t launch_arrival_histogram(const double* arrival, uint32_t n, double lo, double hi, uint32_t n_bins,
                                    unsigned long long* bins, hipStream_t s);
hipError_t launch_multi_step(int integrator, int coll, int precision, const MultiParams& M, const IonPtrs& S,
                             uint32_t block, hipStream_t s);
hipError_t launch_batch(const BatchParams* B, const uint32_t* block, const uint32_t* grid,
                        const uint4* blob, uint32_t n_clouds, unsigned long long* counters, hipStream_t s,
                        const hipStream_t* side, hipEvent_t fork, const hipEvent_t* join);
hipError_t launch_synth_clusters(float4* atoms, const uint32_t* offsets, const uint32_t* counts,
                                 uint32_t n, float r_min, float r_max, uint64_t seed, hipStream_t s);
hipError_t launch_oapa(const float4* atoms, const uint32_t* offsets, const uint32_t* counts, uint32_t n,
                       const double* rot64, double r_gas, double* sigma, hipStream_t s);
hipError_t launch_features(const float4* atoms, const uint32_t* offsets, const uint32_t* counts,
                           uint32_t n, double r_gas, const double* sigma, float* f, hipStream_t s);
hipError_t launch_precompute(const float4* atoms, const uint32_t* offsets, const uint32_t* counts,
                             uint32_t cluster_begin, uint32_t n_clusters, uint32_t n_orient,
                             uint32_t n_sigma, uint32_t n_mu, uint32_t max_tries, uint64_t seed,
                             double r_gas, double* sigma, double* mu, hipStream_t s);
struct TrainBuffers {
    float *X, *T;
    float *Z1, *H1, *Z2, *H2, *Y, *dY, *dH2, *dH1;
    float *params, *grad, *m, *v, *loss;
    float* ws;
    int B;
};
hipError_t mlp_fwd_bwd(const TrainBuffers& t, hipStream_t s);
hipError_t adam_step(const TrainBuffers& t, float lr, float b1, float b2, float eps, int step, hipStream_t s);
hipError_t mlp_infer(const uint4* blob, const float* X, float* Y, uint32_t n, hipStream_t s);
size_t brick_sort_temp_bytes(uint32_t n);
hipError_t random_permutation(uint32_t n, uint64_t seed, uint32_t epoch, uint32_t* keys_a, uint32_t* keys_b,
                              uint32_t* idx, uint32_t* perm, void* temp, size_t temp_bytes, hipStream_t s);
hipError_t sort_ions_by_brick(const double* x, const double* y, const double* z, const uint8_t* flags, uint32_t n,
                              const double origin[3], const double inv_brick[3], const uint32_t nb[3],
                              uint32_t* keys_a, uint32_t* keys_b, uint32_t* idx, uint32_t* order, void* temp,
                              size_t temp_bytes, hipStream_t s);
}

using namespace icarion;

namespace {

thread_local std::string g_last_error;

struct Failure {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw Failure{code, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(ICARION_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return ICARION_OK;
    } catch (const Failure& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return ICARION_ERR_INVALID;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return ICARION_ERR_INVALID;
    }
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void resize(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        hip_check(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)), "hipMalloc");
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

template <class T>
void copy_in(DevBuf<T>& b, const T* src, size_t n, hipStream_t s, const char* what) {
    b.resize(n);
    if (n) hip_check(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s), what);
}

template <class T>
void copy_out(T* dst, const DevBuf<T>& b, size_t n, hipStream_t s, const char* what) {
    if (dst && n) hip_check(hipMemcpyAsync(dst, b.p, n * sizeof(T), hipMemcpyDeviceToHost, s), what);
}

constexpr double kPiHost = 3.14159265358979323846;
constexpr double kBoltzmannHost = 1.380649e-23;

}  // namespace

struct icarion_gpu {
    int device = 0;
    uint32_t block = 256;
    uint32_t steps_per_launch = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    bool have_domain = false;
    icarion_domain_desc dom{};
    icarion_physics_desc phys{};
    StepParams base{};

    bool have_species = false;
    uint32_t n_species = 0;
    std::vector<icarion_species_desc> species_host;
    DevBuf<SpeciesDev> species;
    DevBuf<float4> atoms_f32;
    DevBuf<uint32_t> atom_offsets, atom_counts;   // per species (EHSS tools)
    uint32_t n_atoms = 0;

    bool have_surrogate = false;
    std::vector<float> params;
    DevBuf<uint4> mlp_blob;

    bool have_field = false;
    DevBuf<float4> fa32;
    DevBuf<double4> fa64;
    // the field-array source, kept apart from the domain (set_domain may follow
    // the upload); applied to the kernel parameters at every launch
    struct FieldSrc {
        int nx = 0, ny = 0, nz = 0, n_terms = 0;
        bool nonuniform = false;
        double x0 = 0, y0 = 0, z0 = 0, dx = 1, dy = 1, dz = 1;
        double x_last = 0, y_last = 0, z_last = 0;
        int kind[kMaxFieldTerms] = {0, 0, 0, 0};
        double cst[kMaxFieldTerms] = {0, 0, 0, 0}, freq[kMaxFieldTerms] = {0, 0, 0, 0},
               phase[kMaxFieldTerms] = {0, 0, 0, 0};
    } fsrc;
    DevBuf<double> ax64[3];
    DevBuf<float> ax32[3];
    // field arrays: the ions in grid-brick order for each launch (lane -> ion)
    DevBuf<uint32_t> ord, ord_keys_a, ord_keys_b, ord_idx;
    DevBuf<unsigned char> ord_temp;
    // HSS gas mixture: the participating components with n > 0 (in order) and
    // their per-species cross sections, as the kernels read them
    // voltage waveforms of the analytic fields (per slot; none when !has_waves)
    std::unique_ptr<HostWaveform> wave[ICARION_N_VOLTAGE_SLOTS];
    bool has_waves = false;
    // several domains: each domain's own voltage waveforms (n_dom sets)
    std::vector<std::array<std::unique_ptr<HostWaveform>, ICARION_N_VOLTAGE_SLOTS>> dom_wave;
    // the gas pressure's waveform (EnvironmentConfig::pressure_Pa_waveform;
    // null: the domain's constant pressure), and each domain's on several
    std::unique_ptr<HostWaveform> pressure;
    std::vector<std::unique_ptr<HostWaveform>> dom_pressure;
    std::vector<std::pair<int, double>> dom_flow;   // several domains: (flow model, u_max) each
    // the gas flow profile (enum icarion_flow_model) and its centre-line speed
    int flow_model = ICARION_FLOW_CONSTANT;
    double flow_umax = 0.0;
    // several domains: each domain's HSS gas mixture (filtered components, the
    // cross sections per species, the species count given)
    std::vector<std::vector<icarion_gas_component>> dom_mix;
    std::vector<std::vector<double>> dom_mix_sig;
    std::vector<uint32_t> dom_mix_ns;
    DevBuf<double> dom_mix_d;
    // the same mixtures as given (every component, for the EHSS handler's loop)
    std::vector<std::vector<icarion_gas_component>> dom_mix_all;
    std::vector<std::vector<double>> dom_emix_sig;
    DevBuf<double> dom_emix_d;
    // several domains: each domain's species table (empty: the handle's),
    // its rows concatenated on the device at each step call
    std::vector<std::vector<SpeciesDev>> dom_species;
    DevBuf<SpeciesDev> dom_species_d;
    // several domains: each domain's uniform field array (one unscaled term,
    // float nodes and, for double input, double nodes)
    struct DomField {
        bool have = false;
        int nx = 0, ny = 0, nz = 0;
        double x0 = 0, y0 = 0, z0 = 0, dx = 1, dy = 1, dz = 1;
        DevBuf<float4> fa32;
        DevBuf<double4> fa64;
    };
    std::vector<DomField> dom_field;
    // RK45 evaluates the waveforms in the kernel: their descriptors and points
    DevBuf<icarion_waveform_desc> wave_desc;
    DevBuf<double> wave_pts;
    std::vector<icarion_gas_component> mix;
    uint32_t mix_species = 0;
    DevBuf<double> mix_sigma;
    // the same mixture as given (every component, for the EHSS handler's loop),
    // its full cross-section table and the components' contact radii
    std::vector<icarion_gas_component> mix_all;
    std::vector<double> mix_radius;
    DevBuf<double> emix_sigma;
    // offline EHSS sample sets (per species {first orientation, orientations,
    // mu per orientation, first mu}; none when off_species == 0)
    uint32_t off_species = 0;
    DevBuf<uint4> off_tab;
    DevBuf<double> off_sigma, off_mu;
    // EHSS orientation samples (per species {first orientation, orientations};
    // each orientation's rotation and area; none when ori_species == 0)
    uint32_t ori_species = 0;
    DevBuf<uint2> ori_tab;
    DevBuf<double> ori_R, ori_area;

    // several domains (icarion_gpu_set_domains); n_dom = 0: the single domain
    // above.  Per domain: the descriptor, end aperture^2; the z-sorted spans
    // of the cylindrical domains; each ion's domain index
    uint32_t n_dom = 0;
    std::vector<icarion_domain_desc> doms;
    std::vector<double> dom_ap_sq;
    std::vector<DomainSpan> dom_spans;
    bool dom_has_orbitrap = false;
    DevBuf<StepParams> dom_params_d;
    DevBuf<StepEntry> dom_tables;
    DevBuf<DomainSpan> dom_spans_d;
    DevBuf<double> dom_ap_d;
    DevBuf<int32_t> ion_dom;

    size_t n = 0;
    bool have_state = false;
    DevBuf<double> x, y, z, vx, vy, vz, mass, charge, birth, death, arrival;
    DevBuf<uint32_t> spec, ncoll;
    DevBuf<uint8_t> flags;
    bool has_birth = false;
    uint32_t species_used = 0;   // 1 + the largest species index of the resident ions

    DevBuf<StepEntry> table;
    DevBuf<unsigned long long> counters;
    // the folded counters {ion-steps, collisions, alive, arrived}: mapped
    // pinned host memory and its device address
    unsigned long long* host_ctr = nullptr;
    unsigned long long* host_ctr_dev = nullptr;
    // device address of this device's scratch array (icarion_gpu_step_noalloc)
    unsigned char* scratch_dev = nullptr;
    std::vector<StepEntry> table_host;
    // composite field-array term scales at every stage time (voltage waveforms)
    DevBuf<double> term_scales;
    std::vector<double> term_scales_host;
    bool use_term_scales = false;

    // batched clouds (independent of the single-cloud state above)
    struct BatchCloud {
        size_t n = 0, off = 0;
        int precision = 0;
        int cls = 0;                 // precision class (kBatchClasses)
        uint32_t kind = 0;
        uint64_t n_steps = 0;
        double t_end = 0.0;
        bool has_birth = false;
        bool substeps = false;       // more than one collision stage per step
    };
    uint32_t b_n = 0;
    size_t b_total = 0;
    std::vector<BatchCloud> b_clouds;
    std::vector<StepParams> b_params;
    DevBuf<StepParams> b_params_d;
    DevBuf<IonPtrs> b_ions_d;
    DevBuf<uint32_t> b_kinds_d, b_queue;
    DevBuf<BatchTile> b_tiles[kBatchClasses];
    uint32_t b_n_tiles[kBatchClasses] = {0, 0, 0}, b_tiled_block[kBatchClasses] = {0, 0, 0};
    DevBuf<unsigned long long> b_done, b_start, b_counters;
    DevBuf<double> b_x, b_y, b_z, b_vx, b_vy, b_vz, b_mass, b_charge, b_birth, b_death, b_arrival;
    DevBuf<uint32_t> b_spec, b_ncoll;
    DevBuf<uint8_t> b_flags;
    DevBuf<SpeciesDev> b_species;
    DevBuf<StepEntry> b_tables;
    DevBuf<double> b_mix_sigma;
    DevBuf<float4> b_atoms;
    DevBuf<float4> b_fa32;      // the clouds' field arrays, one copy per distinct array
    icarion_batch_tuning b_tuned[2] = {{0, 0}, {0, 0}};
    hipStream_t side[kBatchClasses - 1] = {nullptr, nullptr};   // streams of the other classes
    hipEvent_t ev_fork = nullptr, ev_join[kBatchClasses - 1] = {nullptr, nullptr};
    int n_cu = 0;
    int clock_khz = 0;

    // trajectory snapshots: kSnapSlots device staging slots, each holding one packed
    // snapshot in full (positions [n][3], velocities [n][3], species [n], domain
    // index [n], flags [n]: 57 B per ion) and the compacted records of the ions that changed
    // since the previous snapshot (56 B each) with their count; the flags at
    // the last snapshot; and the pinned host copy of the current snapshot that
    // the callback reads, updated in full or by the changed records
    struct SnapRing {
        size_t cap = 0;                         // ions per slot
        DevBuf<uint8_t> full[3], rec[3];        // kSnapSlots
        DevBuf<uint32_t> count;                 // [kSnapSlots]
        DevBuf<uint8_t> prev;
        void* host_full = nullptr;              // hipHostMalloc, 57 B per ion
        void* host_rec = nullptr;               // hipHostMalloc, 56 B per ion
        uint32_t* host_count = nullptr;         // hipHostMalloc, [kSnapSlots]
        uint32_t* host_count_dev = nullptr;     // its device address
        hipEvent_t packed[3] = {nullptr, nullptr, nullptr};
        hipEvent_t copied = nullptr;
    } snap;

    // RK45: per-ion time, dt and step-control state; the engine record
    Rk45Cfg rk45{1e-8, 1e-6, 0.9, 1e-6, 1e3, 5.0, 0.1, 0.0, 1, 0};
    DevBuf<double> r_t, r_dt, r_err, r_k1x, r_k1y, r_k1z;
    DevBuf<uint8_t> r_fsal;
    DevBuf<Rk45Engine> r_eng;
};

namespace {

void set_device(const icarion_gpu* h) { hip_check(hipSetDevice(h->device), "hipSetDevice"); }

bool is_identity(const double m[9]) {
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 9; ++i)
        if (m[i] != I[i]) return false;
    return true;
}

int field_kind(const icarion_domain_desc& d) {
    if (d.use_field_array) return FIELD_ARRAY;
    if (d.instrument == ICARION_INSTRUMENT_ORBITRAP) return FIELD_ORBITRAP;
    return FIELD_LINEAR;
}

// The kernel variant of a handle's field: the general field-array path for
// non-uniform axes or several terms.
int launch_field_kind(const icarion_gpu* h) {
    const int f = field_kind(h->dom);
    if (f == FIELD_ARRAY && (h->fsrc.nonuniform || h->fsrc.n_terms > 1)) return FIELD_ARRAY_X;
    return f;
}

int coll_kind(const icarion_physics_desc& p) {
    switch (p.collision_model) {
        case ICARION_COLLISION_HSS: return COLL_HSS;
        case ICARION_COLLISION_EHSS: return COLL_EHSS;
        case ICARION_COLLISION_LANGEVIN:
        case ICARION_COLLISION_FRICTION: return COLL_DAMPING;
        default: return COLL_NONE;
    }
}

// The kernel variant of a handle's collision stage: HSS over a gas mixture
// has its own.
int launch_coll_kind(const icarion_gpu* h) {
    const int c = coll_kind(h->phys);
    if (c == COLL_HSS && !h->mix.empty()) return COLL_HSS_MIX;
    if (c == COLL_EHSS && !h->mix_all.empty() && h->phys.ehss_mode == ICARION_EHSS_GEOMETRY) return COLL_EHSS_MIX;
    if (c == COLL_EHSS && (h->off_species || h->ori_species) && h->phys.ehss_mode == ICARION_EHSS_GEOMETRY)
        return COLL_EHSS_TAB;
    return c;
}

void validate_domain(const icarion_domain_desc& d, const icarion_physics_desc& p) {
    if (d.instrument < 0 || d.instrument > 3) fail(ICARION_ERR_INVALID, "unknown instrument");
    if (p.integrator != ICARION_INTEGRATOR_RK4 && p.integrator != ICARION_INTEGRATOR_BORIS &&
        p.integrator != ICARION_INTEGRATOR_RK45)
        fail(ICARION_ERR_UNSUPPORTED, "integrator not covered by the device path (RK4, Boris, RK45)");
    if (p.collision_model < 0 || p.collision_model > 4) fail(ICARION_ERR_INVALID, "unknown collision model");
    if (p.precision != ICARION_PRECISION_FP64 && p.precision != ICARION_PRECISION_FP32_KAHAN)
        fail(ICARION_ERR_INVALID, "unknown precision");
    if (p.collision_substeps < 0 || p.collision_substeps > ICARION_MAX_COLLISION_SUBSTEPS)
        fail(ICARION_ERR_INVALID, "collision_substeps must be 0 .. ICARION_MAX_COLLISION_SUBSTEPS");
    if (!(d.length_m > 0.0)) fail(ICARION_ERR_INVALID, "geometry.length_m must be > 0");
    if (d.instrument == ICARION_INSTRUMENT_ORBITRAP) {
        if (!(d.radius_in_m > 0.0 && d.radius_out_m > d.radius_in_m && d.radius_char_m > d.radius_out_m))
            fail(ICARION_ERR_INVALID, "Orbitrap needs 0 < radius_in < radius_out < radius_char");
    } else if (!(d.radius_m > 0.0)) {
        fail(ICARION_ERR_INVALID, "geometry.radius_m must be > 0");
    }
    if (p.collision_model != ICARION_COLLISION_NONE) {
        if (!(d.temperature_K > 0.0) || !(d.gas_mass_kg > 0.0))
            fail(ICARION_ERR_INVALID, "collisions need temperature_K > 0 and gas_mass_kg > 0");
    }
}

// The Orbitrap's field constant for a DC radial voltage V
// (AnalyticalFieldModel.cpp:201-202).
double orbitrap_k(const icarion_domain_desc& d, double V) {
    const double r_in = d.radius_in_m, r_out = d.radius_out_m;
    const double rc_sq = d.radius_char_m * d.radius_char_m;
    return 2.0 * V / (rc_sq * std::log(r_out / r_in) - 0.5 * (r_out * r_out - r_in * r_in));
}

// compute_orbitrap_field evaluates the DC radial voltage at t = 0 only
// (AnalyticalFieldModel.cpp:198): a waveform there fixes the constant at its
// t = 0 value, and the Orbitrap reads no other voltage slot.
void apply_orbitrap_waveform(const icarion_gpu* h, StepParams& P) {
    if (h->dom.instrument != ICARION_INSTRUMENT_ORBITRAP || !h->has_waves) return;
    if (const HostWaveform* w = h->wave[ICARION_SLOT_DC_RADIAL].get()) P.orbi_k = orbitrap_k(h->dom, w->eval(0.0));
}

StepParams derive_params(const icarion_domain_desc& d, const icarion_physics_desc& p) {
    StepParams P{};
    P.instrument = d.instrument;
    // the origin shift is always applied; rot_identity skips the 3x3 products
    P.rot_identity = (is_identity(d.rot_g2l) && is_identity(d.rot_l2g)) ? 1 : 0;
    P.magnetic = d.magnetic_enabled;
    P.damping_model = p.collision_model;
    P.ou_kicks = p.ou_thermal_kicks;
    P.n_sub = std::max(1, static_cast<int>(p.collision_substeps));
    P.detect = std::isfinite(d.detector_z_m) ? 1 : 0;
    P.has_rf = (d.instrument != ICARION_INSTRUMENT_ORBITRAP && std::fabs(d.rf_voltage_V) > 1e-12) ? 1 : 0;
    P.rf_U = d.dc_quad_V;
    P.rf_V = d.rf_voltage_V;
    P.rf_omega = 2.0 * kPiHost * d.rf_frequency_Hz;
    P.r0_sq = d.radius_m * d.radius_m;
    for (int i = 0; i < 3; ++i) {
        P.origin[i] = d.origin_m[i];
        P.B[i] = d.B_T[i];
        P.u_gas[i] = d.gas_velocity_m_s[i];
    }
    P.L = d.length_m;
    P.R = d.radius_m;
    P.R_in = d.radius_in_m;
    P.R_out = d.radius_out_m;
    P.R_char = d.radius_char_m;
    P.R_char_sq = d.radius_char_m * d.radius_char_m;
    if (d.instrument == ICARION_INSTRUMENT_ORBITRAP) {
        P.ln_R_in = std::log(d.radius_in_m);
        P.ln_R_out = std::log(d.radius_out_m);
        P.orbi_k = orbitrap_k(d, d.dc_radial_V);
    }
    std::memcpy(P.g2l, d.rot_g2l, sizeof(P.g2l));
    std::memcpy(P.l2g, d.rot_l2g, sizeof(P.l2g));
    if (d.instrument == ICARION_INSTRUMENT_QUADRUPOLE) {
        P.Ez_const = (std::fabs(d.dc_axial_V) > 1e-12) ? d.dc_axial_V / d.length_m : 0.0;
    } else {
        P.Ez_const = d.dc_axial_V / d.length_m;
    }
    P.fa_scale = d.field_array_scale;
    P.T = d.temperature_K;
    P.n = d.gas_density_m3;
    P.m_gas = d.gas_mass_kg;
    P.r_gas = d.gas_radius_m;
    P.alpha_gas = d.gas_polarizability_m3;
    if (d.temperature_K > 0.0 && d.gas_mass_kg > 0.0) {
        P.sigma_neutral = std::sqrt(kBoltzmannHost * d.temperature_K / d.gas_mass_kg);
        P.v_th = std::sqrt(8.0 * kBoltzmannHost * d.temperature_K / (kPiHost * d.gas_mass_kg));
    }
    P.det_z = d.detector_z_m;
    P.seed = p.rng_seed;
    return P;
}

// Per-step table: t_k by repeated addition and c = U_eff(t)/r0^2 at the
// stage times the integrator evaluates (RK4: t, t + dt/2, t + dt; Boris: t).
// Per-step table of a run: t_k by repeated addition from t0 and, with RF,
// c = U_eff(t)/r0^2 at the stage times the integrator evaluates (RK4: t,
// t + dt/2, t + dt; Boris: t).  Appends n_steps entries to out.
// stage_times: a field array of several time-scaled terms, whose kernels take
// the stage times t, t + dt/2, t + dt in place of the RF coefficients.
// pressure: the domain's pressure waveform (may be null).  The engine evaluates
// it at the engine time of each step start (update_dynamic_environments,
// SimulationEngine.cpp:380, SimulationEngineLifecycle.cpp:34-47) and derives
// n = P(t)/(kB T) (EnvironmentConfig.h:132-152); the entry keeps n over the
// domain's constant density, by which the kernels scale every rate.
void append_table(std::vector<StepEntry>& out, const icarion_domain_desc& d, bool rf, int integrator,
                  double t0, double dt, uint64_t n_steps, bool stage_times = false,
                  const HostWaveform* const* wf = nullptr, const HostWaveform* pressure = nullptr) {
    const size_t base = out.size();
    out.resize(base + n_steps);
    const double r0_sq = d.radius_m * d.radius_m;
    const double omega = 2.0 * kPiHost * d.rf_frequency_Hz;
    const double half_dt = dt * 0.5;
    // E_z of the IMS field, or of the quadrupole's axial DC where set
    // (AnalyticalFieldModel.cpp:136-161, :243-245)
    auto ez_of = [&](double axial) {
        if (d.instrument == ICARION_INSTRUMENT_QUADRUPOLE) return std::fabs(axial) > 1e-12 ? axial / d.length_m : 0.0;
        return axial / d.length_m;
    };
    const double ez_const = ez_of(d.dc_axial_V);
    auto coeff = [&](double t) {
        const double U_eff = d.dc_quad_V + d.rf_voltage_V * std::cos(omega * t);
        return U_eff / r0_sq;
    };
    // with waveforms every voltage is evaluated at the stage time, and the RF
    // term is present where |V(t)| > 1e-12 (AnalyticalFieldModel.cpp:150-159, 236-242)
    auto slot = [&](int k, double constant, double t) { return (wf && wf[k]) ? wf[k]->eval(t) : constant; };
    auto coeff_wf = [&](double t) {
        const double V = slot(ICARION_SLOT_RF_VOLTAGE, d.rf_voltage_V, t);
        if (!(std::fabs(V) > 1e-12)) return 0.0;
        const double f = slot(ICARION_SLOT_RF_FREQUENCY, d.rf_frequency_Hz, t);
        const double U = slot(ICARION_SLOT_DC_QUAD, d.dc_quad_V, t);
        const double om = 2.0 * kPiHost * f;
        return (U + V * std::cos(om * t)) / r0_sq;
    };
    auto ez_wf = [&](double t) { return ez_of(slot(ICARION_SLOT_DC_AXIAL, d.dc_axial_V, t)); };
    if (pressure && !(d.gas_density_m3 > 0.0 && d.temperature_K > 0.0))
        fail(ICARION_ERR_INVALID, "a pressure waveform needs the domain's gas density and temperature > 0");
    double t = t0;
    for (uint64_t k = 0; k < n_steps; ++k) {
        StepEntry& e = out[base + k];
        e.t = t;
        e.nrel = 1.0;
        if (pressure) {
            const double p = pressure->eval(t);
            if (!(p > 0.0))   // EnvironmentConfig::update_time_dependent's check (EnvironmentConfig.h:137-139)
                fail(ICARION_ERR_INVALID, "pressure waveform evaluated to a non-positive value at t = " +
                                              std::to_string(t));
            e.nrel = (p / (kBoltzmannHost * d.temperature_K)) / d.gas_density_m3;
        }
        e.e0 = e.eh = e.e1 = ez_const;
        if (stage_times) {
            e.c0 = t;
            e.ch = t + half_dt;
            e.c1 = t + dt;
        } else if (wf) {
            e.c0 = rf ? coeff_wf(t) : 0.0;
            e.e0 = ez_wf(t);
            if (integrator == ICARION_INTEGRATOR_RK4) {
                e.ch = rf ? coeff_wf(t + half_dt) : 0.0;
                e.c1 = rf ? coeff_wf(t + dt) : 0.0;
                e.eh = ez_wf(t + half_dt);
                e.e1 = ez_wf(t + dt);
            } else {
                e.ch = e.c1 = 0.0;
            }
        } else if (rf) {
            e.c0 = coeff(t);
            if (integrator == ICARION_INTEGRATOR_RK4) {
                e.ch = coeff(t + half_dt);
                e.c1 = coeff(t + dt);
            } else {
                e.ch = e.c1 = 0.0;
            }
        } else {
            e.c0 = e.ch = e.c1 = 0.0;
        }
        t = t + dt;
    }
}

// The field-array grid and terms into the launch parameters; each term's
// scale kind resolved against the current domain
// (CompositeFieldProvider::compute_scale_factor, CompositeFieldProvider.h:117-147).
void apply_field(const icarion_gpu* h, StepParams& P) {
    const auto& f = h->fsrc;
    P.nx = f.nx; P.ny = f.ny; P.nz = f.nz;
    P.n_terms = f.n_terms;
    P.x0 = f.x0; P.y0 = f.y0; P.z0 = f.z0;
    P.dx = f.dx; P.dy = f.dy; P.dz = f.dz;
    P.inv_dx = 1.0 / f.dx; P.inv_dy = 1.0 / f.dy; P.inv_dz = 1.0 / f.dz;
    P.x_last = f.x_last; P.y_last = f.y_last; P.z_last = f.z_last;
    P.ax_x = f.nonuniform ? h->ax64[0].p : nullptr;
    P.ax_y = f.nonuniform ? h->ax64[1].p : nullptr;
    P.ax_z = f.nonuniform ? h->ax64[2].p : nullptr;
    P.axf_x = f.nonuniform ? h->ax32[0].p : nullptr;
    P.axf_y = f.nonuniform ? h->ax32[1].p : nullptr;
    P.axf_z = f.nonuniform ? h->ax32[2].p : nullptr;
    const icarion_domain_desc& d = h->dom;
    for (int j = 0; j < kMaxFieldTerms; ++j) {
        P.term_rf[j] = 0;
        P.term_a[j] = 0.0;
        P.term_w[j] = 0.0;
        P.term_phi[j] = 0.0;
        if (j >= f.n_terms) continue;
        switch (f.kind[j]) {
            case ICARION_FIELD_SCALE_CONSTANT: P.term_a[j] = f.cst[j]; break;
            case ICARION_FIELD_SCALE_DC_AXIAL: P.term_a[j] = d.dc_axial_V; break;
            case ICARION_FIELD_SCALE_DC_QUAD: P.term_a[j] = d.dc_quad_V; break;
            case ICARION_FIELD_SCALE_DC_RADIAL: P.term_a[j] = d.dc_radial_V; break;
            default: {
                const double fr = f.freq[j] > 0.0 ? f.freq[j] : d.rf_frequency_Hz;
                P.term_rf[j] = 1;
                P.term_a[j] = d.rf_voltage_V;
                P.term_w[j] = 2.0 * kPiHost * fr;
                P.term_phi[j] = f.phase[j];
            }
        }
    }
}

// The components HSSCollisionHandler's loop reaches (participating, n > 0;
// HSSCollisionHandler.cpp:128-134) and their cross sections per species
// (sig[s * mix.size() + j]); with none, one zero-density component, which
// gives the kernels the handler's k_total = 0.
void build_hss_mixture(const icarion_gas_component* components, uint32_t n_components, const double* sigma_m2,
                       uint32_t n_species, std::vector<icarion_gas_component>& mix, std::vector<double>& sig) {
    if (n_components > static_cast<uint32_t>(kMaxGasComponents))
        fail(ICARION_ERR_UNSUPPORTED, "at most " + std::to_string(kMaxGasComponents) + " gas components");
    std::vector<uint32_t> keep;
    for (uint32_t c = 0; c < n_components; ++c) {
        const icarion_gas_component& g = components[c];
        if (!std::isfinite(g.density_m3) || !std::isfinite(g.mass_kg))
            fail(ICARION_ERR_INVALID, "gas component values must be finite");
        if (!g.participates || g.density_m3 <= 0.0) continue;
        if (!(g.mass_kg > 0.0)) fail(ICARION_ERR_INVALID, "gas component mass must be > 0");
        for (uint32_t sp = 0; sp < n_species; ++sp)
            if (!(sigma_m2[static_cast<size_t>(sp) * n_components + c] > 0.0))
                fail(ICARION_ERR_INVALID, "missing cross section for species " + std::to_string(sp) +
                                              " and gas component " + std::to_string(c));
        keep.push_back(c);
    }
    sig.assign(static_cast<size_t>(n_species) * std::max<size_t>(keep.size(), 1), 0.0);
    for (uint32_t sp = 0; sp < n_species; ++sp)
        for (size_t j = 0; j < keep.size(); ++j)
            sig[sp * keep.size() + j] = sigma_m2[static_cast<size_t>(sp) * n_components + keep[j]];
    mix.clear();
    for (uint32_t c : keep) mix.push_back(components[c]);
    if (mix.empty())
        mix.push_back(icarion_gas_component{0.0, components[0].mass_kg > 0.0 ? components[0].mass_kg : 1.0, 1, 0});
}

// The HSS mixture's per-component values into the launch parameters.
void mixture_params(const std::vector<icarion_gas_component>& mix, double temperature_K, StepParams& P) {
    P.n_mix = static_cast<int>(mix.size());
    for (int c = 0; c < kMaxGasComponents; ++c) {
        const bool on = c < P.n_mix;
        P.mix_n[c] = on ? mix[c].density_m3 : 0.0;
        P.mix_m[c] = on ? mix[c].mass_kg : 0.0;
        // VelocitySampling::sample_neutral_velocity's width (VelocitySampling.cpp:17)
        P.mix_sig[c] = on ? std::sqrt(kBoltzmannHost * temperature_K / mix[c].mass_kg) : 0.0;
    }
}

// The gas mixture into the launch parameters (neutral widths at the domain's
// temperature); an error when the species table changed size since.
void apply_mixture(const icarion_gpu* h, StepParams& P) {
    // a species table replaced after the upload must still cover the ions
    if (h->species_used > h->n_species)
        fail(ICARION_ERR_STATE, "the species table no longer covers the resident ions' species indices");
    P.n_mix = static_cast<int>(h->mix.size());
    P.mix_sigma = h->mix.empty() ? nullptr : h->mix_sigma.p;
    if (!h->mix.empty() && h->mix_species != h->n_species)
        fail(ICARION_ERR_STATE, "the gas mixture's species count differs from the species table");
    P.ehss_off = nullptr;
    P.ehss_off_sigma = nullptr;
    P.ehss_off_mu = nullptr;
    if (h->off_species) {
        if (h->off_species != h->n_species)
            fail(ICARION_ERR_STATE, "the offline EHSS sets' species count differs from the species table");
        P.ehss_off = h->off_tab.p;
        P.ehss_off_sigma = h->off_sigma.p;
        P.ehss_off_mu = h->off_mu.p;
    }
    if (h->ori_species) {
        if (h->ori_species != h->n_species)
            fail(ICARION_ERR_STATE, "the EHSS orientation samples' species count differs from the species table");
        P.ehss_ori = h->ori_tab.p;
        P.ehss_ori_R = h->ori_R.p;
        P.ehss_ori_area = h->ori_area.p;
    }
    mixture_params(h->mix, h->dom.temperature_K, P);
    // EHSS: every component; contact radius = the component's, else the
    // domain's gas radius (EHSSCollisionHandler.cpp:348, 415)
    P.n_emix = static_cast<int>(h->mix_all.size());
    P.emix_sigma = h->mix_all.empty() ? nullptr : h->emix_sigma.p;
    for (int c = 0; c < kMaxGasComponents; ++c) {
        const bool on = c < P.n_emix;
        const icarion_gas_component g = on ? h->mix_all[c] : icarion_gas_component{0.0, 0.0, 0, 0};
        P.emix_part[c] = on && g.participates ? 1 : 0;
        P.emix_n[c] = g.density_m3;
        P.emix_m[c] = g.mass_kg;
        P.emix_sig[c] = (on && g.mass_kg > 0.0) ? std::sqrt(kBoltzmannHost * h->dom.temperature_K / g.mass_kg) : 0.0;
        const double r = (on && c < static_cast<int>(h->mix_radius.size())) ? h->mix_radius[c] : 0.0;
        P.emix_r[c] = r > 0.0 ? r : h->dom.gas_radius_m;
    }
}

// Checks and copies one waveform descriptor.
static std::unique_ptr<HostWaveform> parse_waveform(const icarion_waveform_desc* d) {
    if (d->type < ICARION_WAVEFORM_CONSTANT || d->type > ICARION_WAVEFORM_ARBITRARY)
        fail(ICARION_ERR_INVALID, "unknown waveform type");
    if (d->source != ICARION_WAVEFORM_LIBRARY && d->source != ICARION_WAVEFORM_INLINE)
        fail(ICARION_ERR_INVALID, "unknown waveform source");
    for (double v : d->p)
        if (!std::isfinite(v)) fail(ICARION_ERR_INVALID, "waveform parameters must be finite");
    auto hw = std::make_unique<HostWaveform>();
    hw->d = *d;
    if (d->type == ICARION_WAVEFORM_ARBITRARY) {
        // ArbitraryWaveform::evaluate's own checks (WaveformConfig.cpp:11-17)
        if (!d->times_s || !d->values || d->n_points == 0)
            fail(ICARION_ERR_INVALID, "arbitrary waveform: empty time/value arrays");
        if (d->interp < ICARION_INTERP_LINEAR || d->interp > ICARION_INTERP_CUBIC)
            fail(ICARION_ERR_INVALID, "arbitrary waveform: unknown interpolation");
        hw->ts.assign(d->times_s, d->times_s + d->n_points);
        hw->vs.assign(d->values, d->values + d->n_points);
        for (uint32_t i = 0; i < d->n_points; ++i)
            if (!std::isfinite(hw->ts[i]) || !std::isfinite(hw->vs[i]) ||
                (i && !(hw->ts[i] > hw->ts[i - 1])))
                fail(ICARION_ERR_INVALID, "arbitrary waveform: times must increase, values be finite");
    }
    hw->d.times_s = nullptr;
    hw->d.values = nullptr;
    return hw;
}

// The voltage slots (NULL entries: the constant); true when any slot holds a
// waveform.
static bool parse_waveform_slots(const icarion_waveform_desc* const* slots, std::unique_ptr<HostWaveform>* w) {
    bool any = false;
    for (int k = 0; slots && k < ICARION_N_VOLTAGE_SLOTS; ++k) {
        if (!slots[k]) continue;
        w[k] = parse_waveform(slots[k]);
        any = true;
    }
    return any;
}

// CompositeFieldProvider::eval_voltage (CompositeFieldProvider.h:150-171): the
// domain's constant, else a library waveform at t; an inline waveform gives 0.
double composite_voltage(const icarion_gpu* h, int slot, double constant, double t) {
    const HostWaveform* w = h->wave[slot].get();
    if (!w) return constant;
    if (w->d.source == ICARION_WAVEFORM_INLINE) return 0.0;
    return w->eval(t);
}

// CompositeFieldProvider::compute_scale_factor (CompositeFieldProvider.h:117-147).
double composite_scale(const icarion_gpu* h, int j, double t) {
    const auto& f = h->fsrc;
    const icarion_domain_desc& d = h->dom;
    switch (f.kind[j]) {
        case ICARION_FIELD_SCALE_CONSTANT: return f.cst[j];
        case ICARION_FIELD_SCALE_DC_AXIAL: return composite_voltage(h, ICARION_SLOT_DC_AXIAL, d.dc_axial_V, t);
        case ICARION_FIELD_SCALE_DC_QUAD: return composite_voltage(h, ICARION_SLOT_DC_QUAD, d.dc_quad_V, t);
        case ICARION_FIELD_SCALE_DC_RADIAL: return composite_voltage(h, ICARION_SLOT_DC_RADIAL, d.dc_radial_V, t);
        default: {
            const double V = composite_voltage(h, ICARION_SLOT_RF_VOLTAGE, d.rf_voltage_V, t);
            const double fr = f.freq[j] > 0.0
                                  ? f.freq[j]
                                  : composite_voltage(h, ICARION_SLOT_RF_FREQUENCY, d.rf_frequency_Hz, t);
            const double omega = 2.0 * kPiHost * fr;
            return V * std::cos(omega * t + f.phase[j]);
        }
    }
}

// The call's step table on the host, then on the device: in the handle's table
// (resized to the call), or at dst when the caller gives one (at least n_steps
// entries).
void build_table(icarion_gpu* h, double t0, double dt, uint64_t n_steps, StepEntry* dst = nullptr) {
    h->table_host.clear();
    const bool stage_times = field_kind(h->dom) == FIELD_ARRAY && h->fsrc.n_terms >= 2;
    const HostWaveform* wf[ICARION_N_VOLTAGE_SLOTS];
    for (int k = 0; k < ICARION_N_VOLTAGE_SLOTS; ++k) wf[k] = h->wave[k].get();
    const bool rf = h->base.has_rf != 0 || (h->has_waves && h->wave[ICARION_SLOT_RF_VOLTAGE] &&
                                            h->dom.instrument != ICARION_INSTRUMENT_ORBITRAP);
    append_table(h->table_host, h->dom, rf, h->phys.integrator, t0, dt, n_steps, stage_times,
                 h->has_waves ? wf : nullptr, h->pressure.get());
    // with waveforms, the composite's scales at every stage time go into a
    // table of their own, and the stage values become its row numbers
    h->use_term_scales = stage_times && h->has_waves;
    if (h->use_term_scales) {
        if (3 * n_steps > (1ull << 52)) fail(ICARION_ERR_INVALID, "too many steps for one call");
        h->term_scales_host.assign(3 * n_steps * kMaxFieldTerms, 0.0);
        const double half_dt = dt * 0.5;
        for (uint64_t k = 0; k < n_steps; ++k) {
            StepEntry& e = h->table_host[k];
            const double ts[3] = {e.t, e.t + half_dt, e.t + dt};
            for (int st = 0; st < 3; ++st)
                for (int j = 0; j < h->fsrc.n_terms; ++j)
                    h->term_scales_host[(3 * k + st) * kMaxFieldTerms + j] = composite_scale(h, j, ts[st]);
            e.c0 = static_cast<double>(3 * k);
            e.ch = static_cast<double>(3 * k + 1);
            e.c1 = static_cast<double>(3 * k + 2);
        }
        h->term_scales.resize(h->term_scales_host.size());
        hip_check(hipMemcpyAsync(h->term_scales.p, h->term_scales_host.data(),
                                 h->term_scales_host.size() * sizeof(double), hipMemcpyHostToDevice, h->stream),
                  "upload term scales");
    }
    if (!dst) {
        h->table.resize(std::max<size_t>(n_steps, 1));
        dst = h->table.p;
    }
    hip_check(hipMemcpyAsync(dst, h->table_host.data(), n_steps * sizeof(StepEntry), hipMemcpyHostToDevice,
                             h->stream),
              "upload step table");
}

__global__ void zero_u64(unsigned long long* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

constexpr int kCounterWords = kCounterStripes * kCounterStride;

// Field-array clouds of at least this many ions run in grid-brick order.
constexpr size_t kOrderMinIons = 65536;

// Clear the striped launch counters of a single cloud.
void zero_counters(unsigned long long* counters, hipStream_t s) {
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(256), 0, s, counters, kCounterWords);
    hip_check(hipGetLastError(), "zero counters");
}

// Alive (active and born) and arrived (arrival time set) ions of the resident
// state, counted per wavefront with ballots and added to words 2 and 3 of the
// launch counter stripes.
__global__ void __launch_bounds__(256) count_state_kernel(const uint8_t* __restrict__ flags,
                                                          const double* __restrict__ arrival, uint32_t n,
                                                          unsigned long long* counters, int stripes) {
    uint32_t alive = 0, arrived = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        alive += (flags[i] & 3u) == 3u;
        arrived += arrival[i] == arrival[i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        alive += __shfl_xor(alive, off, 64);
        arrived += __shfl_xor(arrived, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && (alive | arrived)) {
        const uint32_t s =
            ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % static_cast<uint32_t>(stripes)) * kCounterStride;
        if (alive) atomicAdd(&counters[s + 2], static_cast<unsigned long long>(alive));
        if (arrived) atomicAdd(&counters[s + 3], static_cast<unsigned long long>(arrived));
    }
}

// Sum of the stripes: word w of every stripe into out[w] (w = 0..3), written
// straight to mapped pinned host memory.
__global__ void fold_counters_kernel(const unsigned long long* counters, int stripes, unsigned long long* out) {
    const int w = threadIdx.x;
    if (w < 4) {
        unsigned long long a = 0;
        for (int k = 0; k < stripes; ++k) a += counters[k * kCounterStride + w];
        out[w] = a;
    }
}

// Counters of the call: ion-steps and collisions of its launches, and the
// alive / arrived counts of the state they left (stream-ordered; 32 bytes
// cross to the host).
struct CallCounts {
    uint64_t ion_steps = 0, collisions = 0, alive = 0, arrived = 0;
};

CallCounts read_counters(icarion_gpu* h, hipStream_t s, bool with_state, unsigned long long* counters) {
    if (with_state && h->n) {
        const uint32_t n = static_cast<uint32_t>(h->n);
        const uint32_t grid = std::min<uint32_t>((n + 255) / 256, 1024);
        hipLaunchKernelGGL(count_state_kernel, dim3(grid), dim3(256), 0, s, h->flags.p, h->arrival.p, n,
                           counters, kCounterStripes);
        hip_check(hipGetLastError(), "count state launch");
    }
    hipLaunchKernelGGL(fold_counters_kernel, dim3(1), dim3(64), 0, s, counters, kCounterStripes,
                       h->host_ctr_dev);
    hip_check(hipGetLastError(), "fold counters launch");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    CallCounts c;
    c.ion_steps = h->host_ctr[0];
    c.collisions = h->host_ctr[1];
    c.alive = h->host_ctr[2];
    c.arrived = h->host_ctr[3];
    return c;
}

IonPtrs ion_ptrs(icarion_gpu* h) {
    IonPtrs S{};
    S.x = h->x.p; S.y = h->y.p; S.z = h->z.p;
    S.vx = h->vx.p; S.vy = h->vy.p; S.vz = h->vz.p;
    S.mass = h->mass.p; S.charge = h->charge.p;
    S.species = h->spec.p;
    S.flags = h->flags.p;
    S.birth = h->has_birth ? h->birth.p : nullptr;
    S.death = h->death.p;
    S.arrival = h->arrival.p;
    S.ncoll = h->ncoll.p;
    S.domain = h->ion_dom.p;
    return S;
}

// EHSS over a gas mixture: the geometry method only; orientation samples would
// need an area per component gas, which the single-gas tables do not hold;
// every participating component needs a mass for its neutral draw.
void check_ehss_mixture(const icarion_gpu* h, bool surrogate) {
    if (h->mix_all.empty()) return;
    if (surrogate) fail(ICARION_ERR_UNSUPPORTED, "the EHSS surrogate runs with a single gas");
    if (h->ori_species)
        fail(ICARION_ERR_UNSUPPORTED, "EHSS orientation samples hold areas for one gas, not for a mixture");
    for (const auto& g : h->mix_all)
        if (g.participates && !(g.mass_kg > 0.0))
            fail(ICARION_ERR_INVALID, "every participating gas component needs a mass > 0");
}

// RK45 on an analytic IMS / quadrupole field with waveforms (voltages): the
// kernel evaluates every slot at each ion's stage times, from the descriptors
// and ARBITRARY points copied here and each slot's constant.  The pressure
// waveform goes into the descriptor after the voltage slots; the kernel
// evaluates it at each engine step's time.
void upload_rk45_waves(icarion_gpu* h, StepParams& P, bool voltages) {
    const icarion_domain_desc& d = h->dom;
    const double cst[ICARION_N_VOLTAGE_SLOTS] = {d.dc_axial_V, d.dc_quad_V, d.dc_radial_V, d.rf_voltage_V,
                                                 d.rf_frequency_Hz};
    constexpr int kSlots = ICARION_N_VOLTAGE_SLOTS + 1;
    std::vector<icarion_waveform_desc> desc(kSlots);
    std::vector<double> pts;
    size_t off_t[kSlots] = {}, off_v[kSlots] = {};
    uint32_t mask = 0;
    for (int k = 0; k < kSlots; ++k) {
        if (k < ICARION_N_VOLTAGE_SLOTS) P.wave_const[k] = cst[k];
        const HostWaveform* w = k < ICARION_N_VOLTAGE_SLOTS ? (voltages ? h->wave[k].get() : nullptr)
                                                            : h->pressure.get();
        if (!w) continue;
        mask |= 1u << k;
        desc[k] = w->d;
        desc[k].n_points = static_cast<uint32_t>(w->ts.size());
        off_t[k] = pts.size();
        pts.insert(pts.end(), w->ts.begin(), w->ts.end());
        off_v[k] = pts.size();
        pts.insert(pts.end(), w->vs.begin(), w->vs.end());
    }
    h->wave_pts.resize(std::max<size_t>(pts.size(), 1));
    if (!pts.empty()) copy_in(h->wave_pts, pts.data(), pts.size(), h->stream, "waveform points");
    for (int k = 0; k < kSlots; ++k) {
        desc[k].times_s = (mask >> k & 1u) && desc[k].n_points ? h->wave_pts.p + off_t[k] : nullptr;
        desc[k].values = (mask >> k & 1u) && desc[k].n_points ? h->wave_pts.p + off_v[k] : nullptr;
    }
    h->wave_desc.resize(kSlots);
    copy_in(h->wave_desc, desc.data(), desc.size(), h->stream, "waveform descriptors");
    hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    P.waves = h->wave_desc.p;
    P.wave_mask = mask & ((1u << ICARION_N_VOLTAGE_SLOTS) - 1u);
    if (voltages && h->wave[ICARION_SLOT_RF_VOLTAGE]) P.has_rf = 1;
    if (h->pressure) {
        if (!(d.gas_density_m3 > 0.0 && d.temperature_K > 0.0))
            fail(ICARION_ERR_INVALID, "a pressure waveform needs the domain's gas density and temperature > 0");
        P.pressure_wave = 1;
        P.pressure_scale = 1.0 / (kBoltzmannHost * d.temperature_K * d.gas_density_m3);
    }
}

// The RK45 engine loop: one launch pair (step kernel + engine-time advance) per
// engine step, replayed from a captured HIP graph of up to 32 steps.
void do_step_rk45(icarion_gpu* h, double t0, double dt, uint64_t n_steps, uint64_t step_offset,
                  icarion_step_stats* stats) {
    const int field = launch_field_kind(h);
    const int coll = launch_coll_kind(h);
    if (h->has_waves && field != FIELD_ORBITRAP && field != FIELD_LINEAR)
        fail(ICARION_ERR_UNSUPPORTED, "RK45 on a field array runs with constant voltages");
    if (h->phys.precision == ICARION_PRECISION_FP32_KAHAN && (coll == COLL_EHSS || coll == COLL_EHSS_TAB))
        fail(ICARION_ERR_UNSUPPORTED, "EHSS geometry collisions run in FP64 only");
    if (coll == COLL_EHSS && h->phys.ehss_mode == ICARION_EHSS_SURROGATE)
        fail(ICARION_ERR_UNSUPPORTED, "the EHSS surrogate runs with the fixed-step FP32 path");
    if (coll == COLL_EHSS) check_ehss_mixture(h, false);
    set_device(h);
    hipStream_t s = h->stream;
    StepParams P = h->base;
    P.dt = dt;
    P.species = h->species.p;
    P.atoms = h->atoms_f32.p;
    P.fa_f32 = h->fa32.p;
    P.fa_f64 = h->fa64.p;
    apply_field(h, P);
    apply_mixture(h, P);
    P.n_ions = static_cast<uint32_t>(h->n);
    P.counters = h->counters.p;
    P.counter_stripes = kCounterStripes;
    P.rk45 = h->rk45;
    apply_orbitrap_waveform(h, P);
    if ((h->has_waves && field == FIELD_LINEAR) || h->pressure)
        upload_rk45_waves(h, P, h->has_waves && field == FIELD_LINEAR);
    if (h->flow_model == ICARION_FLOW_AXIAL_PARABOLIC) {
        P.flow_parabolic = 1;
        P.flow_umax = h->flow_umax;
        P.flow_R2 = h->dom.radius_m > 0.0 ? h->dom.radius_m * h->dom.radius_m : 0.0;
    }
    fill_fast_const(P);
    const IonPtrs S = ion_ptrs(h);
    h->r_eng.resize(1);
    Rk45Ptrs R{h->r_t.p, h->r_dt.p, h->r_err.p, h->r_k1x.p, h->r_k1y.p, h->r_k1z.p, h->r_fsal.p, h->r_eng.p};
    Rk45Engine e0{t0, dt, step_offset, 0ull, ~0ull, 0ull, 0ull, 0ull};
    hip_check(hipMemcpyAsync(h->r_eng.p, &e0, sizeof(e0), hipMemcpyHostToDevice, s), "upload engine record");
    zero_counters(h->counters.p, s);
    float ms = 0.0f;
    if (h->n > 0) {
        hip_check(launch_rk45_init(R, static_cast<uint32_t>(h->n), t0, dt, s), "rk45 init");
        hip_check(hipEventRecord(h->ev0, s), "hipEventRecord");
        const uint64_t G = std::min<uint64_t>(n_steps, 32);
        uint64_t done = 0;
        if (n_steps >= 2 * G && G > 1) {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            hipError_t ce = hipSuccess;
            for (uint64_t k = 0; k < G && ce == hipSuccess; ++k)
                ce = launch_rk45_step(field, coll, h->phys.precision, P, S, R, h->block, s);
            const hipError_t ee = hipStreamEndCapture(s, &graph);
            if (ce == hipErrorInvalidValue) {
                if (graph) (void)hipGraphDestroy(graph);
                fail(ICARION_ERR_UNSUPPORTED, "no RK45 kernel for this field/collision");
            }
            hip_check(ce, "capture RK45 steps");
            hip_check(ee, "hipStreamEndCapture");
            const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(graph);
                hip_check(ie, "hipGraphInstantiate");
            }
            hipError_t le = hipSuccess;
            for (; done + G <= n_steps && le == hipSuccess; done += G) le = hipGraphLaunch(exec, s);
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            hip_check(le, "hipGraphLaunch");
        }
        for (; done < n_steps; ++done) {
            const hipError_t e = launch_rk45_step(field, coll, h->phys.precision, P, S, R, h->block, s);
            if (e == hipErrorInvalidValue) fail(ICARION_ERR_UNSUPPORTED, "no RK45 kernel for this field/collision");
            hip_check(e, "RK45 step launch");
        }
        hip_check(hipEventRecord(h->ev1, s), "hipEventRecord");
        hip_check(hipEventSynchronize(h->ev1), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(&ms, h->ev0, h->ev1), "hipEventElapsedTime");
    } else {
        // no ions: the engine time still advances by the fallback dt
        Rk45Engine e_empty = e0;
        for (uint64_t k = 0; k < n_steps; ++k) e_empty.T = e_empty.T + dt;
        e_empty.step = step_offset + n_steps;
        hip_check(hipMemcpyAsync(h->r_eng.p, &e_empty, sizeof(e_empty), hipMemcpyHostToDevice, s),
                  "upload engine record");
    }
    Rk45Engine e1{};
    hip_check(hipMemcpyAsync(&e1, h->r_eng.p, sizeof(e1), hipMemcpyDeviceToHost, s), "read engine record");
    const CallCounts c = read_counters(h, s, stats != nullptr, h->counters.p);
    if (e1.bad_pressure)   // the steps before it ran, as the reference's engine does before it throws
        fail(ICARION_ERR_INVALID, "pressure waveform evaluated to a non-positive value at t = " + std::to_string(e1.T));
    if (e1.failed)
        fail(ICARION_ERR_INVALID, "RK45: " + std::to_string(e1.failed) +
                                      " ion steps found no accepted step in 10 attempts");
    if (stats) {
        stats->ion_steps = c.ion_steps;
        stats->collisions = c.collisions;
        stats->rejected = e1.rejected;
        stats->t_end_s = e1.T;
        stats->device_ms = ms;
        stats->alive = c.alive;
        stats->arrived = c.arrived;
    }
}

// Trajectory sampling of one icarion_gpu_step_snapshots call.
struct SnapshotReq {
    double interval;
    bool first;
    double* next;
    icarion_snapshot_fn fn;
    void* user;
};

// One planned snapshot: after local step k, at engine time t; next = the next
// write time once it is recorded.
struct SnapshotAt {
    uint64_t k;
    double t, next;
};

// The engine's sampling rule (SimulationEngine.cpp:275-278): after its first
// step, and after every step whose end time reaches the next write time, which
// then becomes that time + the interval (OutputManager.cpp:230, 246-253).  The
// step end times are the table's start times + dt, the engine's accumulation.
std::vector<SnapshotAt> plan_snapshots(const std::vector<StepEntry>& table, double dt, uint64_t n_steps,
                                       const SnapshotReq& r) {
    std::vector<SnapshotAt> out;
    double next = *r.next;
    for (uint64_t k = 0; k < n_steps; ++k) {
        const double t = table[k].t + dt;
        if ((k == 0 && r.first) || t >= next) {
            next = t + r.interval;
            out.push_back({k, t, next});
        }
    }
    return out;
}

// Snapshots in flight: snapshot q is brought to the host once snapshot q + 2 is
// queued, so two write intervals of steps keep the device busy meanwhile.
constexpr int kSnapSlots = 3;
constexpr size_t kSnapFullBytes = 57;     // per ion: 6 doubles, species, domain, flags
constexpr size_t kSnapRecordBytes = 56;   // per changed ion: 6 doubles, index, flags, domain

void ensure_snap_ring(icarion_gpu* h) {
    auto& R = h->snap;
    // a call that ended in an error may have left a pack or a copy queued on
    // the ring's buffers: let them finish before the buffers are reused or freed
    hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    hip_check(hipStreamSynchronize(h->side[0]), "hipStreamSynchronize");
    if (!R.copied) hip_check(hipEventCreateWithFlags(&R.copied, hipEventDisableTiming), "hipEventCreate");
    for (int j = 0; j < kSnapSlots; ++j)
        if (!R.packed[j]) hip_check(hipEventCreateWithFlags(&R.packed[j], hipEventDisableTiming), "hipEventCreate");
    if (!R.host_count) {
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&R.host_count), kSnapSlots * sizeof(uint32_t),
                                hipHostMallocMapped | hipHostMallocCoherent),
                  "hipHostMalloc snapshot count");
        hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&R.host_count_dev), R.host_count, 0),
                  "hipHostGetDevicePointer");
        R.count.resize(kSnapSlots);
    }
    if (R.cap >= h->n && R.host_full) return;
    R.cap = 0;
    if (R.host_full) (void)hipHostFree(R.host_full);
    if (R.host_rec) (void)hipHostFree(R.host_rec);
    R.host_full = R.host_rec = nullptr;
    for (int j = 0; j < kSnapSlots; ++j) {
        R.full[j].resize(kSnapFullBytes * h->n);
        R.rec[j].resize(kSnapRecordBytes * h->n);
    }
    R.prev.resize(h->n);
    hip_check(hipHostMalloc(&R.host_full, kSnapFullBytes * h->n, hipHostMallocDefault), "hipHostMalloc snapshot");
    hip_check(hipHostMalloc(&R.host_rec, kSnapRecordBytes * h->n, hipHostMallocDefault), "hipHostMalloc snapshot");
    R.cap = h->n;
}

// The launch loop of one stepping call: launches of at most spl steps, cut at
// the snapshot steps of plan (empty: no snapshots).  After each snapshot step
// the state is packed on the step stream and moved to the host while the next
// steps run.  launch(k0, nk) queues steps k0 .. k0 + nk - 1 of the call.
// Returns the device time from the first launch to the last (ms).
template <class Launch>
float run_launches(icarion_gpu* h, const IonPtrs& S, uint64_t n_steps, uint64_t step_offset, uint64_t spl,
                   const std::vector<SnapshotAt>& plan, const SnapshotReq* snap, Launch&& launch) {
    auto& R = h->snap;
    const size_t n = h->n;
    size_t q_issued = 0, q_done = 0;
    bool stopped = false;
    uint64_t stop_step = 0;
    // the pinned snapshot the callback sees (none when the call takes no snapshots)
    uint8_t* const hf = plan.empty() ? nullptr : static_cast<uint8_t*>(R.host_full);
    double* const hpos = hf ? reinterpret_cast<double*>(hf) : nullptr;
    double* const hvel = hf ? hpos + 3 * n : nullptr;
    uint32_t* const hspec = hf ? reinterpret_cast<uint32_t*>(hf + 48 * n) : nullptr;
    int32_t* const hdom = hf ? reinterpret_cast<int32_t*>(hf + 52 * n) : nullptr;
    uint8_t* const hflags = hf ? hf + 56 * n : nullptr;
    // Snapshot q_done moves to the host in two stages, each started as soon
    // as the one before has finished, so the calling thread keeps queueing
    // steps meanwhile: once its pack (and count) is done, the copy starts
    // on a side stream - of the changed ions' records when they are few,
    // else of the whole snapshot; once the copy is done, the records are
    // scattered over the previous snapshot and fn is called.  block = wait
    // (spinning) until snapshot q_done is delivered.
    bool copying = false, compact = false;
    size_t n_rec = 0;
    auto ready = [](hipEvent_t e, bool block) {
        for (;;) {
            const hipError_t r = hipEventQuery(e);
            if (r == hipSuccess) return true;
            if (r != hipErrorNotReady) hip_check(r, "hipEventQuery");
            if (!block) return false;
        }
    };
    auto advance = [&](bool block) {
        while (q_done < q_issued) {
            const int j = static_cast<int>(q_done % kSnapSlots);
            if (!copying) {
                if (!ready(R.packed[j], block)) return;
                n_rec = R.host_count[j];
                if (n_rec > n) fail(ICARION_ERR_HIP, "snapshot record count exceeds the ion count");
                compact = q_done > 0 && n_rec * 8 <= n;
                if (!compact)
                    hip_check(hipMemcpyAsync(hf, R.full[j].p, kSnapFullBytes * n, hipMemcpyDeviceToHost,
                                             h->side[0]),
                              "snapshot copy");
                else if (n_rec)
                    hip_check(hipMemcpyAsync(R.host_rec, R.rec[j].p, kSnapRecordBytes * n_rec,
                                             hipMemcpyDeviceToHost, h->side[0]),
                              "snapshot record copy");
                hip_check(hipEventRecord(R.copied, h->side[0]), "hipEventRecord");
                copying = true;
            }
            if (!ready(R.copied, block)) return;
            copying = false;
            if (compact) {
                const uint8_t* r = static_cast<const uint8_t*>(R.host_rec);
                for (size_t k = 0; k < n_rec; ++k, r += kSnapRecordBytes) {
                    double pv[6];
                    uint32_t idx;
                    std::memcpy(pv, r, sizeof(pv));
                    std::memcpy(&idx, r + 48, sizeof(idx));
                    if (idx >= n) fail(ICARION_ERR_HIP, "snapshot record index out of range");
                    std::memcpy(hpos + 3 * size_t(idx), pv, 3 * sizeof(double));
                    std::memcpy(hvel + 3 * size_t(idx), pv + 3, 3 * sizeof(double));
                    hflags[idx] = r[52];
                    hdom[idx] = static_cast<int8_t>(r[53]);
                }
            }
            const SnapshotAt& a = plan[q_done];
            *snap->next = a.next;
            ++q_done;
            if (snap->fn(snap->user, q_done - 1, step_offset + a.k, a.t, n, hpos, hvel, hspec, hdom, hflags) != 0) {
                stopped = true;
                stop_step = step_offset + a.k;
                return;
            }
            block = false;
        }
    };
    // Pack snapshot q_issued on the step stream, its count published to
    // pinned memory behind it; first free its slot if every slot is in flight.
    auto issue = [&]() {
        if (q_issued - q_done == kSnapSlots) advance(true);
        if (stopped) return;
        const int j = static_cast<int>(q_issued % kSnapSlots);
        uint8_t* d = R.full[j].p;
        double* pos = reinterpret_cast<double*>(d);
        if (q_issued == 0) hip_check(launch_zero_u32(R.count.p, h->stream), "snapshot count clear");
        hip_check(launch_pack_snapshot(S, static_cast<uint32_t>(n), q_issued == 0 ? 1 : 0, pos, pos + 3 * n,
                                       reinterpret_cast<uint32_t*>(d + 48 * n),
                                       reinterpret_cast<int32_t*>(d + 52 * n), d + 56 * n, R.prev.p,
                                       R.rec[j].p, R.count.p + j, R.host_count_dev + j,
                                       R.count.p + (j + 1) % kSnapSlots, h->stream),
                  "snapshot pack launch");
        hip_check(hipEventRecord(R.packed[j], h->stream), "hipEventRecord");
        ++q_issued;
        advance(false);
    };
    hip_check(hipEventRecord(h->ev0, h->stream), "hipEventRecord");
    size_t q_next = 0;
    uint64_t k0 = 0;
    while (k0 < n_steps && !stopped) {
        uint64_t k1 = std::min<uint64_t>(k0 + spl, n_steps);
        if (q_next < plan.size()) k1 = std::min<uint64_t>(k1, plan[q_next].k + 1);
        launch(k0, k1 - k0);
        if (q_next < plan.size() && plan[q_next].k + 1 == k1) {
            issue();
            ++q_next;
        } else if (q_done < q_issued) {
            advance(false);
        }
        k0 = k1;
    }
    hip_check(hipEventRecord(h->ev1, h->stream), "hipEventRecord");
    while (q_done < q_issued && !stopped) advance(true);
    if (copying) (void)hipStreamSynchronize(h->side[0]);
    float ms = 0.0f;
    hip_check(hipEventSynchronize(h->ev1), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(&ms, h->ev0, h->ev1), "hipEventElapsedTime");
    if (stopped)
        fail(ICARION_ERR_STATE, "snapshot callback ended the run after step " + std::to_string(stop_step) +
                                    "; steps through " + std::to_string(step_offset + k0 - 1) + " ran");
    return ms;
}

void do_step_multi(icarion_gpu* h, double t0, double dt, uint64_t n_steps, uint64_t step_offset,
                   icarion_step_stats* stats, const SnapshotReq* snap);

// A snapshot schedule over a call with no resident ions: the callback still
// sees every due snapshot, with n = 0 and no arrays.
void deliver_empty_snapshots(const icarion_domain_desc& d, int integrator, double t0, double dt, uint64_t n_steps,
                             uint64_t step_offset, const SnapshotReq& snap) {
    std::vector<StepEntry> tab;
    append_table(tab, d, false, integrator, t0, dt, n_steps);
    const std::vector<SnapshotAt> plan = plan_snapshots(tab, dt, n_steps, snap);
    for (size_t q = 0; q < plan.size(); ++q) {
        *snap.next = plan[q].next;
        if (snap.fn(snap.user, q, step_offset + plan[q].k, plan[q].t, 0, nullptr, nullptr, nullptr, nullptr,
                    nullptr) != 0)
            fail(ICARION_ERR_STATE, "snapshot callback ended the run after step " +
                                        std::to_string(step_offset + plan[q].k));
    }
}

void finish_stats(icarion_gpu* h, double t_end, float ms, icarion_step_stats* stats, unsigned long long* counters);

// Device memory of a fixed-step call besides the ion state: its step table
// and launch counters.  By default the handle's own buffers (the table is
// resized to the call); a caller that may not allocate passes fixed memory
// that holds up to max_steps table entries.
struct Workspace {
    StepEntry* table = nullptr;
    unsigned long long* counters = nullptr;
    uint64_t max_steps = 0;
};

void do_step(icarion_gpu* h, double t0, double dt, uint64_t n_steps, uint64_t step_offset,
             icarion_step_stats* stats, const SnapshotReq* snap = nullptr, const Workspace* ws = nullptr) {
    if (ws) {
        // the allocation-free variant covers the single-domain fixed-step path
        if (h->have_domain && h->n_dom >= 2)
            fail(ICARION_ERR_UNSUPPORTED, "the allocation-free step covers single-domain handles");
        if (h->have_domain && h->phys.integrator == ICARION_INTEGRATOR_RK45)
            fail(ICARION_ERR_UNSUPPORTED, "the allocation-free step covers the fixed-step integrators");
        if (snap) fail(ICARION_ERR_UNSUPPORTED, "the allocation-free step takes no snapshots");
        if (n_steps > ws->max_steps)
            fail(ICARION_ERR_INVALID, "n_steps " + std::to_string(n_steps) + " exceeds the allocation-free "
                                      "step's maximum of " + std::to_string(ws->max_steps) + " per call");
        if (h->have_domain && h->has_waves && field_kind(h->dom) == FIELD_ARRAY && h->fsrc.n_terms >= 2)
            fail(ICARION_ERR_UNSUPPORTED, "waveform-scaled field-array terms need a per-call scale table; "
                                          "use icarion_gpu_step");
    }
    unsigned long long* const counters = ws ? ws->counters : h->counters.p;
    if (h->have_domain && h->n_dom >= 2) {
        do_step_multi(h, t0, dt, n_steps, step_offset, stats, snap);
        return;
    }
    if (h->have_domain && h->phys.integrator == ICARION_INTEGRATOR_RK45) {
        if (snap) fail(ICARION_ERR_UNSUPPORTED, "snapshots are taken on the fixed-step integrators");
        if (!h->have_species) fail(ICARION_ERR_STATE, "icarion_gpu_set_species not called");
        if (!h->have_state) fail(ICARION_ERR_STATE, "icarion_gpu_upload_state not called");
        if (!(dt > 0.0) || !std::isfinite(dt)) fail(ICARION_ERR_INVALID, "dt must be finite and > 0");
        if (!std::isfinite(t0)) fail(ICARION_ERR_INVALID, "t0 must be finite");
        if (field_kind(h->dom) == FIELD_ARRAY && !h->have_field)
            fail(ICARION_ERR_STATE, "domain uses a field array but none was uploaded");
        do_step_rk45(h, t0, dt, n_steps, step_offset, stats);
        return;
    }
    if (!h->have_domain) fail(ICARION_ERR_STATE, "icarion_gpu_set_domain not called");
    if (!h->have_species) fail(ICARION_ERR_STATE, "icarion_gpu_set_species not called");
    if (!h->have_state) fail(ICARION_ERR_STATE, "icarion_gpu_upload_state not called");
    if (!(dt > 0.0) || !std::isfinite(dt)) fail(ICARION_ERR_INVALID, "dt must be finite and > 0");
    if (!std::isfinite(t0)) fail(ICARION_ERR_INVALID, "t0 must be finite");
    const int field = field_kind(h->dom);
    if (field == FIELD_ARRAY && !h->have_field)
        fail(ICARION_ERR_STATE, "domain uses a field array but none was uploaded");
    const int coll = coll_kind(h->phys);
    const bool surrogate = coll == COLL_EHSS && h->phys.ehss_mode == ICARION_EHSS_SURROGATE;
    if (surrogate && h->phys.precision != ICARION_PRECISION_FP32_KAHAN)
        fail(ICARION_ERR_UNSUPPORTED, "the EHSS surrogate runs in the FP32 + Kahan path");
    if (surrogate && !h->have_surrogate)
        fail(ICARION_ERR_STATE, "EHSS surrogate selected but no parameters set");
    if (coll == COLL_EHSS && !surrogate && h->phys.precision == ICARION_PRECISION_FP32_KAHAN)
        fail(ICARION_ERR_UNSUPPORTED, "EHSS geometry collisions run in FP64 only");
    if (coll == COLL_EHSS) check_ehss_mixture(h, surrogate);
    // the wave-cooperative geometry EHSS kernel and the surrogate kernel run
    // one collision stage per step
    if (h->base.n_sub > 1 && (surrogate || launch_coll_kind(h) == COLL_EHSS))
        fail(ICARION_ERR_UNSUPPORTED, "collision substeps with EHSS run with offline sets, orientation "
                                      "samples, gas mixtures, several domains, RK45 or the batch");
    const bool parabolic = h->flow_model == ICARION_FLOW_AXIAL_PARABOLIC;
    if (parabolic && (surrogate || launch_coll_kind(h) == COLL_EHSS))
        fail(ICARION_ERR_UNSUPPORTED, "the parabolic gas flow runs with HSS, damping, or EHSS with offline sets, "
                                      "orientation samples or a gas mixture");
    set_device(h);
    StepParams P = h->base;
    if (parabolic) {
        P.flow_parabolic = 1;
        P.flow_umax = h->flow_umax;
        P.flow_R2 = h->dom.radius_m > 0.0 ? h->dom.radius_m * h->dom.radius_m : 0.0;
    }
    if (h->has_waves && h->wave[ICARION_SLOT_RF_VOLTAGE]) P.has_rf = 1;
    P.dt = dt;
    P.species = h->species.p;
    P.atoms = h->atoms_f32.p;
    P.mlp_blob = h->mlp_blob.p;
    P.fa_f32 = h->fa32.p;
    P.fa_f64 = (h->phys.precision == ICARION_PRECISION_FP64) ? h->fa64.p : nullptr;
    apply_field(h, P);
    apply_mixture(h, P);
    P.n_ions = static_cast<uint32_t>(h->n);
    P.counters = counters;
    P.counter_stripes = kCounterStripes;
    apply_orbitrap_waveform(h, P);
    fill_fast_const(P);
    zero_counters(counters, h->stream);
    double t_end = t0;
    float ms = 0.0f;
    if (snap && n_steps > 0 && h->n == 0)
        deliver_empty_snapshots(h->dom, h->phys.integrator, t0, dt, n_steps, step_offset, *snap);
    if (n_steps > 0 && h->n > 0) {
        build_table(h, t0, dt, n_steps, ws ? ws->table : nullptr);
        StepEntry* const table = ws ? ws->table : h->table.p;
        P.term_scales = h->use_term_scales ? h->term_scales.p : nullptr;
        t_end = h->table_host[n_steps - 1].t + dt;
        const uint64_t spl = h->steps_per_launch ? h->steps_per_launch : 2000;
        const IonPtrs S = ion_ptrs(h);
        std::vector<SnapshotAt> plan;
        if (snap) {
            plan = plan_snapshots(h->table_host, dt, n_steps, *snap);
            ensure_snap_ring(h);
        }
        // field arrays: each launch runs the ions in grid-brick order (bricks of
        // 4 x 4 x 4 nodes), sorted on the device from the state the launch
        // starts from; small clouds and the allocation-free path keep lane = ion
        const bool order = !ws && field == FIELD_ARRAY && h->n >= kOrderMinIons;
        double o_origin[3] = {0, 0, 0}, o_inv[3] = {0, 0, 0};
        uint32_t o_nb[3] = {1, 1, 1};
        if (order) {
            const uint32_t n = static_cast<uint32_t>(h->n);
            h->ord.resize(n);
            h->ord_keys_a.resize(n);
            h->ord_keys_b.resize(n);
            h->ord_idx.resize(n);
            const size_t tb = brick_sort_temp_bytes(n);
            if (h->ord_temp.n < tb) h->ord_temp.resize(tb);
            const auto& f = h->fsrc;
            const double first[3] = {f.x0, f.y0, f.z0}, last[3] = {f.x_last, f.y_last, f.z_last};
            const uint32_t nn[3] = {static_cast<uint32_t>(f.nx), static_cast<uint32_t>(f.ny), static_cast<uint32_t>(f.nz)};
            for (int a = 0; a < 3; ++a) {
                const double spacing = nn[a] > 1 ? (last[a] - first[a]) / (nn[a] - 1) : 1.0;
                o_origin[a] = first[a];
                o_nb[a] = std::max<uint32_t>(1, (nn[a] + 2) / 4);
                o_inv[a] = spacing > 0.0 ? 1.0 / (4.0 * spacing) : 0.0;
            }
        }
        ms = run_launches(h, S, n_steps, step_offset, spl, plan, snap, [&](uint64_t k0, uint64_t nk) {
            P.table = table + k0;
            P.step0 = step_offset + k0;
            P.n_steps = static_cast<uint32_t>(nk);
            P.order = nullptr;
            if (order) {
                hip_check(sort_ions_by_brick(h->x.p, h->y.p, h->z.p, h->flags.p, static_cast<uint32_t>(h->n),
                                             o_origin, o_inv, o_nb, h->ord_keys_a.p, h->ord_keys_b.p, h->ord_idx.p,
                                             h->ord.p, h->ord_temp.p, h->ord_temp.n, h->stream),
                          "brick sort");
                P.order = h->ord.p;
            }
            const hipError_t e = launch_step(h->phys.integrator, launch_field_kind(h), launch_coll_kind(h),
                                             h->phys.precision, surrogate ? 1 : 0, P, S, h->block, h->stream);
            if (e == hipErrorInvalidValue)
                fail(ICARION_ERR_UNSUPPORTED, "no kernel for this integrator/field/collision/precision");
            hip_check(e, "step kernel launch");
        });
    } else {
        for (uint64_t k = 0; k < n_steps; ++k) t_end = t_end + dt;
    }
    finish_stats(h, t_end, ms, stats, counters);
}


// Alive / arrived counts and the counters into stats (stream-ordered reads;
// the counts are reduced on the device).
void finish_stats(icarion_gpu* h, double t_end, float ms, icarion_step_stats* stats, unsigned long long* counters) {
    if (!stats) {
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        return;
    }
    const CallCounts c = read_counters(h, h->stream, true, counters);
    stats->ion_steps = c.ion_steps;
    stats->collisions = c.collisions;
    stats->rejected = 0;
    stats->t_end_s = t_end;
    stats->device_ms = ms;
    stats->alive = c.alive;
    stats->arrived = c.arrived;
}

// The engine loop over several domains: per-domain parameter records and
// coefficient tables, the span list, then one multi-domain launch per
// steps_per_launch steps.
void do_step_multi(icarion_gpu* h, double t0, double dt, uint64_t n_steps, uint64_t step_offset,
                   icarion_step_stats* stats, const SnapshotReq* snap) {
    if (!h->have_species) fail(ICARION_ERR_STATE, "icarion_gpu_set_species not called");
    if (!h->have_state) fail(ICARION_ERR_STATE, "icarion_gpu_upload_state not called");
    if (!(dt > 0.0) || !std::isfinite(dt)) fail(ICARION_ERR_INVALID, "dt must be finite and > 0");
    if (!std::isfinite(t0)) fail(ICARION_ERR_INVALID, "t0 must be finite");
    if (h->phys.integrator != ICARION_INTEGRATOR_RK4 && h->phys.integrator != ICARION_INTEGRATOR_BORIS)
        fail(ICARION_ERR_UNSUPPORTED, "several domains run with RK4 or Boris");

    const int coll = coll_kind(h->phys);
    if (coll == COLL_EHSS) {
        if (h->phys.ehss_mode == ICARION_EHSS_SURROGATE)
            fail(ICARION_ERR_UNSUPPORTED, "several domains: the EHSS surrogate is not covered");
        if (h->phys.precision != ICARION_PRECISION_FP64)
            fail(ICARION_ERR_UNSUPPORTED, "EHSS geometry collisions run in FP64 only");
        if (h->off_species || h->ori_species)
            fail(ICARION_ERR_UNSUPPORTED, "several domains: offline EHSS sets and orientation samples are not covered");
    }
    if (!h->mix.empty())
        fail(ICARION_ERR_UNSUPPORTED, "several domains take their gas mixtures per domain (icarion_gpu_set_domain_gas_mixture)");
    bool dom_mixture = false;
    for (uint32_t d = 0; d < h->n_dom; ++d)
        if (!h->dom_species[d].empty() && h->dom_species[d].size() != h->n_species)
            fail(ICARION_ERR_STATE, "a domain's species table differs in size from the species table");
    for (uint32_t d = 0; d < h->n_dom; ++d) {
        if (h->dom_mix[d].empty()) continue;
        dom_mixture = true;
        if (h->dom_mix_ns[d] != h->n_species)
            fail(ICARION_ERR_STATE, "a domain's gas mixture species count differs from the species table");
    }
    if (dom_mixture && coll != COLL_HSS && coll != COLL_EHSS)
        fail(ICARION_ERR_UNSUPPORTED, "per-domain gas mixtures run with HSS or geometry EHSS collisions");
    if (h->has_waves)
        fail(ICARION_ERR_UNSUPPORTED, "several domains take their waveforms per domain (icarion_gpu_set_domain_waveforms)");
    if (h->pressure)
        fail(ICARION_ERR_UNSUPPORTED, "several domains take their pressure waveforms per domain "
                                      "(icarion_gpu_set_domain_pressure_waveform)");
    if (h->flow_model != ICARION_FLOW_CONSTANT)
        fail(ICARION_ERR_UNSUPPORTED, "several domains take their gas flows per domain (icarion_gpu_set_domain_gas_flow)");
    bool dom_fields = false;
    bool any_flow = false;
    for (uint32_t d = 0; d < h->n_dom; ++d) {
        if (!h->doms[d].use_field_array) continue;
        if (!h->dom_field[d].have)
            fail(ICARION_ERR_STATE, "domain " + std::to_string(d) +
                                        " uses a field array but none was uploaded (icarion_gpu_upload_domain_field_array)");
        dom_fields = true;
    }
    if (h->species_used > h->n_species)
        fail(ICARION_ERR_STATE, "the species table no longer covers the resident ions' species indices");
    set_device(h);
    const uint32_t nd = h->n_dom;
    zero_counters(h->counters.p, h->stream);
    double t_end = t0;
    float ms = 0.0f;
    if (snap && n_steps > 0 && h->n == 0)
        deliver_empty_snapshots(h->doms[0], h->phys.integrator, t0, dt, n_steps, step_offset, *snap);
    if (n_steps > 0 && h->n > 0) {
        std::vector<StepParams> Ps(nd);
        // the domains' mixture cross sections, concatenated
        std::vector<double> msig;
        std::vector<size_t> moff(nd, 0);
        for (uint32_t d = 0; d < nd; ++d) {
            moff[d] = msig.size();
            msig.insert(msig.end(), h->dom_mix_sig[d].begin(), h->dom_mix_sig[d].end());
        }
        h->dom_mix_d.resize(std::max<size_t>(msig.size(), 1));
        if (!msig.empty()) copy_in(h->dom_mix_d, msig.data(), msig.size(), h->stream, "domain mixture cross sections");
        std::vector<double> esig;
        std::vector<size_t> eoff(nd, 0);
        for (uint32_t d = 0; d < nd; ++d) {
            eoff[d] = esig.size();
            esig.insert(esig.end(), h->dom_emix_sig[d].begin(), h->dom_emix_sig[d].end());
        }
        h->dom_emix_d.resize(std::max<size_t>(esig.size(), 1));
        if (!esig.empty()) copy_in(h->dom_emix_d, esig.data(), esig.size(), h->stream, "domain mixture cross sections");
        // the domains' own species tables, concatenated
        std::vector<SpeciesDev> dsp;
        std::vector<size_t> soff(nd, 0);
        for (uint32_t d = 0; d < nd; ++d) {
            soff[d] = dsp.size();
            dsp.insert(dsp.end(), h->dom_species[d].begin(), h->dom_species[d].end());
        }
        h->dom_species_d.resize(std::max<size_t>(dsp.size(), 1));
        if (!dsp.empty()) copy_in(h->dom_species_d, dsp.data(), dsp.size(), h->stream, "domain species tables");
        std::vector<StepEntry> tabs;
        tabs.reserve(size_t(nd) * n_steps);
        for (uint32_t d = 0; d < nd; ++d) {
            StepParams& P = Ps[d];
            P = derive_params(h->doms[d], h->phys);
            P.dt = dt;
            P.species = h->dom_species[d].empty() ? h->species.p : h->dom_species_d.p + soff[d];
            P.atoms = h->atoms_f32.p;
            P.n_ions = static_cast<uint32_t>(h->n);
            P.n_mix = 0;
            P.mix_sigma = nullptr;
            if (!h->dom_mix[d].empty()) {
                mixture_params(h->dom_mix[d], h->doms[d].temperature_K, P);
                P.mix_sigma = h->dom_mix_d.p + moff[d];
            }
            // EHSS: every component, contact radius the domain's gas radius
            // (EHSSCollisionHandler.cpp:311-424)
            const auto& all = h->dom_mix_all[d];
            P.n_emix = static_cast<int>(all.size());
            P.emix_sigma = all.empty() ? nullptr : h->dom_emix_d.p + eoff[d];
            for (int c = 0; c < kMaxGasComponents; ++c) {
                const bool on = c < P.n_emix;
                const icarion_gas_component g = on ? all[c] : icarion_gas_component{0.0, 0.0, 0, 0};
                P.emix_part[c] = on && g.participates ? 1 : 0;
                P.emix_n[c] = g.density_m3;
                P.emix_m[c] = g.mass_kg;
                P.emix_sig[c] =
                    (on && g.mass_kg > 0.0) ? std::sqrt(kBoltzmannHost * h->doms[d].temperature_K / g.mass_kg) : 0.0;
                P.emix_r[c] = h->doms[d].gas_radius_m;
            }
            // the domain's own waveforms, as a single domain's (AnalyticalFieldModel.cpp:130-248)
            const HostWaveform* wf[ICARION_N_VOLTAGE_SLOTS];
            bool waves = false;
            for (int k = 0; k < ICARION_N_VOLTAGE_SLOTS; ++k) {
                wf[k] = h->dom_wave[d][k].get();
                waves = waves || wf[k];
            }
            if (wf[ICARION_SLOT_RF_VOLTAGE]) P.has_rf = 1;
            if (h->doms[d].instrument == ICARION_INSTRUMENT_ORBITRAP && wf[ICARION_SLOT_DC_RADIAL])
                P.orbi_k = orbitrap_k(h->doms[d], wf[ICARION_SLOT_DC_RADIAL]->eval(0.0));
            P.field_array = 0;
            if (h->doms[d].use_field_array) {
                // GridFieldProvider: E from the domain's own array, unscaled for
                // one term (PhysicsSetup.cpp:168-171) apart from field_array_scale
                const auto& F = h->dom_field[d];
                P.field_array = 1;
                P.nx = F.nx; P.ny = F.ny; P.nz = F.nz;
                P.n_terms = 1;
                P.x0 = F.x0; P.y0 = F.y0; P.z0 = F.z0;
                P.dx = F.dx; P.dy = F.dy; P.dz = F.dz;
                P.inv_dx = 1.0 / F.dx; P.inv_dy = 1.0 / F.dy; P.inv_dz = 1.0 / F.dz;
                P.x_last = F.x0 + (F.nx - 1) * F.dx;
                P.y_last = F.y0 + (F.ny - 1) * F.dy;
                P.z_last = F.z0 + (F.nz - 1) * F.dz;
                P.ax_x = P.ax_y = P.ax_z = nullptr;
                P.axf_x = P.axf_y = P.axf_z = nullptr;
                for (int j = 0; j < kMaxFieldTerms; ++j) {
                    P.term_rf[j] = 0;
                    P.term_a[j] = P.term_w[j] = P.term_phi[j] = 0.0;
                }
                P.fa_f32 = F.fa32.p;
                P.fa_f64 = h->phys.precision == ICARION_PRECISION_FP64 ? F.fa64.p : nullptr;
            }
            // the domain's AxialParabolic gas flow (EnvironmentConfig::gas_velocity_at)
            if (h->dom_flow[d].first == ICARION_FLOW_AXIAL_PARABOLIC) {
                P.flow_parabolic = 1;
                P.flow_umax = h->dom_flow[d].second;
                P.flow_R2 = h->doms[d].radius_m > 0.0 ? h->doms[d].radius_m * h->doms[d].radius_m : 0.0;
                any_flow = true;
            }
            fill_fast_const(P);
            append_table(tabs, h->doms[d], P.has_rf != 0, h->phys.integrator, t0, dt, n_steps, false,
                         waves ? wf : nullptr, h->dom_pressure[d].get());
        }
        t_end = tabs[n_steps - 1].t + dt;
        h->dom_tables.resize(tabs.size());
        copy_in(h->dom_tables, tabs.data(), tabs.size(), h->stream, "domain tables");
        for (uint32_t d = 0; d < nd; ++d) Ps[d].table = h->dom_tables.p + size_t(d) * n_steps;
        h->dom_params_d.resize(nd);
        copy_in(h->dom_params_d, Ps.data(), nd, h->stream, "domain parameters");
        h->dom_ap_d.resize(nd);
        copy_in(h->dom_ap_d, h->dom_ap_sq.data(), nd, h->stream, "domain apertures");
        h->dom_spans_d.resize(std::max<size_t>(h->dom_spans.size(), 1));
        if (!h->dom_spans.empty())
            copy_in(h->dom_spans_d, h->dom_spans.data(), h->dom_spans.size(), h->stream, "domain spans");
        MultiParams M{};
        M.dom = h->dom_params_d.p;
        M.spans = h->dom_spans_d.p;
        M.t_table = h->dom_tables.p;
        M.ap_sq = h->dom_ap_d.p;
        M.species = h->species.p;
        M.counters = h->counters.p;
        M.n_dom = static_cast<int>(nd);
        M.n_spans = static_cast<int>(h->dom_spans.size());
        M.has_orbitrap = h->dom_has_orbitrap ? 1 : 0;
        M.counter_stripes = kCounterStripes;
        M.handoff_tol = 1e-9;   // shared_boundary_handoff_domain's default (DomainManager.h:194-197)
        M.dt = dt;
        M.seed = h->phys.rng_seed;
        M.n_ions = static_cast<uint32_t>(h->n);
        M.n_sub = Ps[0].n_sub;
        M.has_field_array = dom_fields ? 1 : 0;
        M.has_flow = any_flow ? 1 : 0;
        const IonPtrs S = ion_ptrs(h);
        const uint64_t spl = h->steps_per_launch ? h->steps_per_launch : 2000;
        const int integ = h->phys.integrator == ICARION_INTEGRATOR_BORIS ? 1 : 0;
        std::vector<SnapshotAt> plan;
        if (snap) {
            // the engine times are the same in every domain's table
            plan = plan_snapshots(std::vector<StepEntry>(tabs.begin(), tabs.begin() + n_steps), dt, n_steps, *snap);
            ensure_snap_ring(h);
        }
        ms = run_launches(h, S, n_steps, step_offset, spl, plan, snap, [&](uint64_t k0, uint64_t nk) {
            M.k0 = static_cast<uint32_t>(k0);
            M.step0 = step_offset + k0;
            M.n_steps = static_cast<uint32_t>(nk);
            const int kcoll = !dom_mixture ? coll : (coll == COLL_EHSS ? COLL_EHSS_MIX : COLL_HSS_MIX);
            const hipError_t e = launch_multi_step(integ, kcoll, h->phys.precision, M, S, h->block, h->stream);
            if (e == hipErrorInvalidValue) fail(ICARION_ERR_UNSUPPORTED, "no multi-domain kernel for this case");
            hip_check(e, "multi-domain step launch");
        });
    } else {
        for (uint64_t k = 0; k < n_steps; ++k) t_end = t_end + dt;
    }
    finish_stats(h, t_end, ms, stats, h->counters.p);
}

void do_upload(icarion_gpu* h, const icarion_ion_soa* ions) {
    if (!ions) fail(ICARION_ERR_INVALID, "ions is NULL");
    if (!h->have_species) fail(ICARION_ERR_STATE, "icarion_gpu_set_species must precede upload_state");
    const size_t n = ions->n;
    if (n > 0xFFFFFFFFull) fail(ICARION_ERR_INVALID, "at most 2^32-1 ions per handle");
    if (n && (!ions->x || !ions->y || !ions->z || !ions->vx || !ions->vy || !ions->vz ||
              !ions->mass_kg || !ions->charge_C || !ions->species || !ions->active))
        fail(ICARION_ERR_INVALID, "x, y, z, vx, vy, vz, mass_kg, charge_C, species, active are required");
    uint32_t used = 0;
    for (size_t i = 0; i < n; ++i) {
        if (ions->species[i] >= h->n_species)
            fail(ICARION_ERR_INVALID, "species index " + std::to_string(ions->species[i]) +
                                          " of ion " + std::to_string(i) + " out of range");
        if (!(ions->mass_kg[i] > 0.0))
            fail(ICARION_ERR_INVALID, "mass_kg of ion " + std::to_string(i) + " must be > 0");
        used = std::max(used, ions->species[i] + 1);
    }
    h->species_used = used;
    set_device(h);
    hipStream_t s = h->stream;
    copy_in(h->x, ions->x, n, s, "upload x");
    copy_in(h->y, ions->y, n, s, "upload y");
    copy_in(h->z, ions->z, n, s, "upload z");
    copy_in(h->vx, ions->vx, n, s, "upload vx");
    copy_in(h->vy, ions->vy, n, s, "upload vy");
    copy_in(h->vz, ions->vz, n, s, "upload vz");
    copy_in(h->mass, ions->mass_kg, n, s, "upload mass");
    copy_in(h->charge, ions->charge_C, n, s, "upload charge");
    copy_in(h->spec, ions->species, n, s, "upload species");
    std::vector<uint8_t> fl(n);
    for (size_t i = 0; i < n; ++i)
        fl[i] = static_cast<uint8_t>((ions->active[i] ? 1 : 0) | ((ions->born ? ions->born[i] : 1) ? 2 : 0));
    copy_in(h->flags, fl.data(), n, s, "upload flags");
    h->has_birth = ions->birth_time_s != nullptr;
    if (h->has_birth) copy_in(h->birth, ions->birth_time_s, n, s, "upload birth");
    // host staging buffers stay alive until the stream synchronisation below
    std::vector<double> dead(ions->death_time_s ? 0 : n, -1.0);
    std::vector<double> arr(ions->arrival_time_s ? 0 : n, NAN);
    copy_in(h->death, ions->death_time_s ? ions->death_time_s : dead.data(), n, s, "upload death");
    copy_in(h->arrival, ions->arrival_time_s ? ions->arrival_time_s : arr.data(), n, s, "upload arrival");
    std::vector<uint32_t> nc(n, 0);
    copy_in(h->ncoll, ions->n_collisions ? ions->n_collisions : nc.data(), n, s, "upload ncoll");
    // RK45 state: ion time / dt from the caller (NaN / <= 0 = set by the first
    // step), fresh step-control memory (RK45Strategy.h:192-198)
    std::vector<double> tt(ions->time_s ? 0 : n, NAN), dd(ions->dt_s ? 0 : n, 0.0);
    std::vector<double> ones(n, 1.0), zeros(n, 0.0);
    std::vector<uint8_t> nofsal(n, 0);
    copy_in(h->r_t, ions->time_s ? ions->time_s : tt.data(), n, s, "upload time");
    copy_in(h->r_dt, ions->dt_s ? ions->dt_s : dd.data(), n, s, "upload dt");
    copy_in(h->r_err, ones.data(), n, s, "upload rk45 error");
    copy_in(h->r_k1x, zeros.data(), n, s, "upload rk45 k1");
    copy_in(h->r_k1y, zeros.data(), n, s, "upload rk45 k1");
    copy_in(h->r_k1z, zeros.data(), n, s, "upload rk45 k1");
    copy_in(h->r_fsal, nofsal.data(), n, s, "upload rk45 fsal");
    // domain index before the first step: 0, IonState's default (IonState.h:48);
    // several domains set it at each step start
    std::vector<int32_t> dom0(n, 0);
    copy_in(h->ion_dom, dom0.data(), n, s, "upload domain index");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    h->n = n;
    h->have_state = true;
}

void do_download(icarion_gpu* h, icarion_ion_soa* ions) {
    if (!ions) fail(ICARION_ERR_INVALID, "ions is NULL");
    if (!h->have_state) fail(ICARION_ERR_STATE, "no resident state");
    if (ions->n != h->n) fail(ICARION_ERR_INVALID, "ions->n differs from the resident ensemble size");
    set_device(h);
    const size_t n = h->n;
    hipStream_t s = h->stream;
    copy_out(ions->x, h->x, n, s, "download x");
    copy_out(ions->y, h->y, n, s, "download y");
    copy_out(ions->z, h->z, n, s, "download z");
    copy_out(ions->vx, h->vx, n, s, "download vx");
    copy_out(ions->vy, h->vy, n, s, "download vy");
    copy_out(ions->vz, h->vz, n, s, "download vz");
    copy_out(ions->death_time_s, h->death, n, s, "download death");
    copy_out(ions->arrival_time_s, h->arrival, n, s, "download arrival");
    copy_out(ions->n_collisions, h->ncoll, n, s, "download ncoll");
    copy_out(ions->time_s, h->r_t, n, s, "download time");
    copy_out(ions->dt_s, h->r_dt, n, s, "download dt");
    std::vector<uint8_t> fl(n);
    copy_out(fl.data(), h->flags, n, s, "download flags");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    for (size_t i = 0; i < n; ++i) {
        if (ions->active) ions->active[i] = fl[i] & 1;
        if (ions->born) ions->born[i] = (fl[i] >> 1) & 1;
    }
}

// ---- batched clouds --------------------------------------------------------

// kinds from here on (gas mixtures 12-15, geometry EHSS 16-19, field arrays
// 20-25) are in their own kernel variant
constexpr uint32_t kBatchFirstMixKind = 12;

uint32_t batch_kind(const icarion_physics_desc& p, int field, int coll) {
    // the surrogate class has its own kernel: integrator + 2 * Orbitrap
    const uint32_t integ = p.integrator == ICARION_INTEGRATOR_BORIS ? 1u : 0u;
    if (field == FIELD_ARRAY)
        return 20u + integ + 2u * (coll == COLL_HSS ? 1u : coll == COLL_DAMPING ? 2u : 0u);
    const uint32_t c = coll == COLL_HSS       ? 1u
                       : coll == COLL_DAMPING ? 2u
                       : coll == COLL_HSS_MIX ? 3u
                       : coll == COLL_EHSS    ? 4u
                                              : 0u;
    return integ + 2u * (field == FIELD_ORBITRAP ? 1u : 0u) + 4u * c;
}

void do_batch_upload(icarion_gpu* h, const icarion_cloud_desc* clouds, uint32_t n_clouds) {
    if (!clouds || n_clouds == 0) fail(ICARION_ERR_INVALID, "no clouds");
    std::vector<icarion_gpu::BatchCloud> bc(n_clouds);
    std::vector<StepParams> params(n_clouds);
    std::vector<uint32_t> kinds(n_clouds);
    std::vector<SpeciesDev> species;
    std::vector<uint32_t> species_off(n_clouds);
    std::vector<StepEntry> tables;
    std::vector<size_t> table_off(n_clouds);
    std::vector<double> mix_sigma;           // the clouds' mixture cross sections
    std::vector<size_t> mix_off(n_clouds);
    std::vector<float4> atoms_all;           // the clouds' EHSS atom tables
    std::vector<size_t> atom_off(n_clouds);
    std::vector<float4> fa_all;              // the clouds' field arrays (x, y, z, 0)
    std::vector<size_t> fa_off(n_clouds, SIZE_MAX);
    struct FieldKey {
        const float* E;
        uint32_t n[3];
        size_t off;
    };
    std::vector<FieldKey> fields_seen;
    // identical step tables are shared: (t0, dt, n, RF inputs, axial field, stage layout)
    // E_z rides in the table too (e0 / eh / e1), so it is part of the key;
    // clouds with waveforms get tables of their own
    struct Key {
        double t0, dt, U, V, f, r0, Ez;
        uint64_t n;
        int rf, rk4;
        bool operator==(const Key& o) const {
            return std::memcmp(this, &o, sizeof(Key)) == 0;
        }
    };
    std::vector<std::pair<Key, size_t>> seen;
    size_t total = 0;
    for (uint32_t c = 0; c < n_clouds; ++c) {
        const icarion_cloud_desc& cd = clouds[c];
        const std::string tag = "cloud " + std::to_string(c) + ": ";
        validate_domain(cd.domain, cd.physics);
        if (cd.physics.integrator == ICARION_INTEGRATOR_RK45)
            fail(ICARION_ERR_UNSUPPORTED, tag + "RK45 is not covered by the batch");
        const int coll = coll_kind(cd.physics);
        const int field = field_kind(cd.domain);
        const bool surrogate = coll == COLL_EHSS && cd.physics.ehss_mode == ICARION_EHSS_SURROGATE;
        const bool ehss_geo = coll == COLL_EHSS && !surrogate;
        if (ehss_geo && cd.physics.precision != ICARION_PRECISION_FP64)
            fail(ICARION_ERR_UNSUPPORTED, tag + "EHSS geometry collisions run in FP64 only");
        if (cd.n_atoms && !cd.atoms_xyzr) fail(ICARION_ERR_INVALID, tag + "atoms_xyzr is NULL");
        if (surrogate && cd.physics.precision != ICARION_PRECISION_FP32_KAHAN)
            fail(ICARION_ERR_UNSUPPORTED, tag + "the EHSS surrogate runs in the FP32 + Kahan path");
        if (surrogate && cd.physics.collision_substeps > 1)
            fail(ICARION_ERR_UNSUPPORTED, tag + "the EHSS surrogate runs one collision stage per step");
        const bool mixture = cd.n_components > 0;
        if (field == FIELD_ARRAY) {
            if (!cd.field_E) fail(ICARION_ERR_INVALID, tag + "the domain uses a field array but field_E is NULL");
            if (cd.field_n[0] < 2 || cd.field_n[1] < 2 || cd.field_n[2] < 2)
                fail(ICARION_ERR_INVALID, tag + "field array needs >= 2 nodes per axis");
            for (int a = 0; a < 3; ++a)
                if (!(cd.field_spacing_m[a] > 0.0) || !std::isfinite(cd.field_origin_m[a]))
                    fail(ICARION_ERR_INVALID, tag + "field grid spacing must be > 0 and the origin finite");
            if (mixture || coll == COLL_EHSS)
                fail(ICARION_ERR_UNSUPPORTED, tag + "a field array is covered by the batch without gas "
                                                    "mixtures and EHSS");
        }
        if (mixture && (!cd.mixture || !cd.mixture_sigma_m2)) fail(ICARION_ERR_INVALID, tag + "mixture arrays are NULL");
        if (mixture && coll != COLL_HSS)
            fail(ICARION_ERR_UNSUPPORTED, tag + "a gas mixture is covered by the batch with HSS collisions");
        if (!(cd.dt_s > 0.0) || !std::isfinite(cd.dt_s) || !std::isfinite(cd.t0_s))
            fail(ICARION_ERR_INVALID, tag + "t0 must be finite and dt finite and > 0");
        if (cd.n_steps >= 0xFFFFFFFFull) fail(ICARION_ERR_INVALID, tag + "n_steps must be < 2^32");
        if (!cd.species || cd.n_species == 0) fail(ICARION_ERR_INVALID, tag + "species table is empty");
        const icarion_ion_soa* io = cd.ions;
        if (!io) fail(ICARION_ERR_INVALID, tag + "ions is NULL");
        const size_t n = io->n;
        if (n && (!io->x || !io->y || !io->z || !io->vx || !io->vy || !io->vz || !io->mass_kg ||
                  !io->charge_C || !io->species || !io->active))
            fail(ICARION_ERR_INVALID, tag + "x, y, z, vx, vy, vz, mass_kg, charge_C, species, active are required");
        for (size_t i = 0; i < n; ++i) {
            if (io->species[i] >= cd.n_species) fail(ICARION_ERR_INVALID, tag + "species index out of range");
            if (!(io->mass_kg[i] > 0.0)) fail(ICARION_ERR_INVALID, tag + "mass_kg must be > 0");
        }
        bc[c].n = n;
        bc[c].off = total;
        bc[c].precision = cd.physics.precision;
        bc[c].cls = surrogate ? 2 : (cd.physics.precision == ICARION_PRECISION_FP32_KAHAN ? 1 : 0);
        bc[c].substeps = cd.physics.collision_substeps > 1;
        bc[c].kind = surrogate ? static_cast<uint32_t>(cd.physics.integrator == ICARION_INTEGRATOR_BORIS ? 1 : 0) +
                                     2u * (field == FIELD_ORBITRAP ? 1u : 0u)
                               : batch_kind(cd.physics, field, mixture ? COLL_HSS_MIX : coll);
        atom_off[c] = atoms_all.size();
        for (uint32_t a = 0; a < cd.n_atoms; ++a)
            atoms_all.push_back(float4{cd.atoms_xyzr[4 * a], cd.atoms_xyzr[4 * a + 1], cd.atoms_xyzr[4 * a + 2],
                                       cd.atoms_xyzr[4 * a + 3]});
        bc[c].n_steps = cd.n_steps;
        bc[c].has_birth = io->birth_time_s != nullptr;
        total += n;
        species_off[c] = static_cast<uint32_t>(species.size());
        for (uint32_t k = 0; k < cd.n_species; ++k) {
            const icarion_species_desc& d = cd.species[k];
            SpeciesDev sd{};
            sd.sigma_hss = d.sigma_hss_m2;
            sd.sigma_ehss = d.sigma_ehss_m2;
            sd.K0_cm2_Vs = d.K0_cm2_Vs;
            for (int f = 0; f < 3; ++f) sd.feat[f] = d.surrogate_features[f];
            if (ehss_geo) {
                if (static_cast<uint64_t>(d.atom_offset) + d.atom_count > cd.n_atoms)
                    fail(ICARION_ERR_INVALID, tag + "species " + std::to_string(k) + " atoms exceed the atom table");
                sd.atom_offset = d.atom_offset;
                sd.atom_count = d.atom_count;
            }
            species.push_back(sd);
        }
        StepParams P = derive_params(cd.domain, cd.physics);
        P.order = nullptr;
        if (field == FIELD_ARRAY) {
            // one uniform grid, one term, as icarion_gpu_upload_field_array
            const uint32_t* nn = cd.field_n;
            size_t off = SIZE_MAX;
            for (const auto& f : fields_seen)
                if (f.E == cd.field_E && f.n[0] == nn[0] && f.n[1] == nn[1] && f.n[2] == nn[2]) off = f.off;
            if (off == SIZE_MAX) {
                off = fa_all.size();
                const size_t nodes = static_cast<size_t>(nn[0]) * nn[1] * nn[2];
                fa_all.resize(off + nodes);
                for (size_t i = 0; i < nodes; ++i)
                    fa_all[off + i] = make_float4(cd.field_E[3 * i], cd.field_E[3 * i + 1], cd.field_E[3 * i + 2], 0.0f);
                fields_seen.push_back({cd.field_E, {nn[0], nn[1], nn[2]}, off});
            }
            fa_off[c] = off;
            P.nx = static_cast<int>(nn[0]); P.ny = static_cast<int>(nn[1]); P.nz = static_cast<int>(nn[2]);
            P.n_terms = 1;
            P.x0 = cd.field_origin_m[0]; P.y0 = cd.field_origin_m[1]; P.z0 = cd.field_origin_m[2];
            P.dx = cd.field_spacing_m[0]; P.dy = cd.field_spacing_m[1]; P.dz = cd.field_spacing_m[2];
            P.inv_dx = 1.0 / P.dx; P.inv_dy = 1.0 / P.dy; P.inv_dz = 1.0 / P.dz;
            P.x_last = P.x0 + (nn[0] - 1) * P.dx;
            P.y_last = P.y0 + (nn[1] - 1) * P.dy;
            P.z_last = P.z0 + (nn[2] - 1) * P.dz;
            P.ax_x = P.ax_y = P.ax_z = nullptr;
            P.axf_x = P.axf_y = P.axf_z = nullptr;
            for (int j = 0; j < kMaxFieldTerms; ++j) {
                P.term_rf[j] = 0;
                P.term_a[j] = P.term_w[j] = P.term_phi[j] = 0.0;
            }
        }
        mix_off[c] = SIZE_MAX;
        if (mixture) {
            std::vector<icarion_gas_component> mix;
            std::vector<double> sig;
            build_hss_mixture(cd.mixture, cd.n_components, cd.mixture_sigma_m2, cd.n_species, mix, sig);
            mixture_params(mix, cd.domain.temperature_K, P);
            mix_off[c] = mix_sigma.size();
            mix_sigma.insert(mix_sigma.end(), sig.begin(), sig.end());
        }
        // the cloud's voltage waveforms, as icarion_gpu_set_waveforms on a handle of its own
        std::unique_ptr<HostWaveform> wv[ICARION_N_VOLTAGE_SLOTS];
        const bool waves = parse_waveform_slots(cd.waveforms, wv);
        // the cloud's pressure waveform (its step table of its own, as with voltages)
        const std::unique_ptr<HostWaveform> pw = cd.pressure ? parse_waveform(cd.pressure) : nullptr;
        const bool own_table = waves || pw;
        const HostWaveform* wf[ICARION_N_VOLTAGE_SLOTS];
        for (int k = 0; k < ICARION_N_VOLTAGE_SLOTS; ++k) wf[k] = wv[k].get();
        if (wf[ICARION_SLOT_RF_VOLTAGE]) P.has_rf = 1;
        if (field == FIELD_ORBITRAP && wf[ICARION_SLOT_DC_RADIAL])
            P.orbi_k = orbitrap_k(cd.domain, wf[ICARION_SLOT_DC_RADIAL]->eval(0.0));
        Key key{};
        key.t0 = cd.t0_s;
        key.dt = cd.dt_s;
        key.U = cd.domain.dc_quad_V;
        key.V = cd.domain.rf_voltage_V;
        key.f = cd.domain.rf_frequency_Hz;
        key.r0 = cd.domain.radius_m;
        key.Ez = P.Ez_const;
        key.n = cd.n_steps;
        key.rf = P.has_rf;
        key.rk4 = cd.physics.integrator == ICARION_INTEGRATOR_RK4;
        size_t off = SIZE_MAX;
        if (!own_table)
            for (const auto& kv : seen)
                if (kv.first == key) off = kv.second;
        if (off == SIZE_MAX) {
            off = tables.size();
            append_table(tables, cd.domain, P.has_rf != 0, cd.physics.integrator, cd.t0_s, cd.dt_s, cd.n_steps,
                         false, waves ? wf : nullptr, pw.get());
            if (!own_table) seen.push_back({key, off});
        }
        table_off[c] = off;
        bc[c].t_end = cd.n_steps ? tables[off + cd.n_steps - 1].t + cd.dt_s : cd.t0_s;
        P.dt = cd.dt_s;
        P.step0 = cd.step_offset;
        P.n_steps = static_cast<uint32_t>(cd.n_steps);
        P.n_ions = static_cast<uint32_t>(n);
        fill_fast_const(P);
        params[c] = P;
        kinds[c] = bc[c].kind;
    }
    if (total > 0xFFFFFFFFull) fail(ICARION_ERR_INVALID, "at most 2^32-1 ions per batch");
    set_device(h);
    hipStream_t s = h->stream;
    // concatenated SoA; host staging lives until the synchronisation below
    std::vector<double> hx(total), hy(total), hz(total), hvx(total), hvy(total), hvz(total), hm(total),
        hq(total), hb(total, -INFINITY), hd(total, -1.0), ha(total, NAN);
    std::vector<uint32_t> hs(total), hn(total, 0);
    std::vector<uint8_t> hf(total);
    for (uint32_t c = 0; c < n_clouds; ++c) {
        const icarion_ion_soa* io = clouds[c].ions;
        const size_t o = bc[c].off;
        for (size_t i = 0; i < bc[c].n; ++i) {
            hx[o + i] = io->x[i]; hy[o + i] = io->y[i]; hz[o + i] = io->z[i];
            hvx[o + i] = io->vx[i]; hvy[o + i] = io->vy[i]; hvz[o + i] = io->vz[i];
            hm[o + i] = io->mass_kg[i]; hq[o + i] = io->charge_C[i];
            hs[o + i] = io->species[i];
            hf[o + i] = static_cast<uint8_t>((io->active[i] ? 1 : 0) | ((io->born ? io->born[i] : 1) ? 2 : 0));
            if (io->birth_time_s) hb[o + i] = io->birth_time_s[i];
            if (io->death_time_s) hd[o + i] = io->death_time_s[i];
            if (io->arrival_time_s) ha[o + i] = io->arrival_time_s[i];
            if (io->n_collisions) hn[o + i] = io->n_collisions[i];
        }
    }
    const size_t nt = std::max<size_t>(total, 1);
    h->b_x.resize(nt); h->b_y.resize(nt); h->b_z.resize(nt);
    copy_in(h->b_x, hx.data(), total, s, "batch x"); copy_in(h->b_y, hy.data(), total, s, "batch y");
    copy_in(h->b_z, hz.data(), total, s, "batch z"); copy_in(h->b_vx, hvx.data(), total, s, "batch vx");
    copy_in(h->b_vy, hvy.data(), total, s, "batch vy"); copy_in(h->b_vz, hvz.data(), total, s, "batch vz");
    copy_in(h->b_mass, hm.data(), total, s, "batch mass"); copy_in(h->b_charge, hq.data(), total, s, "batch charge");
    copy_in(h->b_birth, hb.data(), total, s, "batch birth"); copy_in(h->b_death, hd.data(), total, s, "batch death");
    copy_in(h->b_arrival, ha.data(), total, s, "batch arrival"); copy_in(h->b_spec, hs.data(), total, s, "batch species");
    copy_in(h->b_ncoll, hn.data(), total, s, "batch ncoll"); copy_in(h->b_flags, hf.data(), total, s, "batch flags");
    copy_in(h->b_species, species.data(), species.size(), s, "batch species table");
    copy_in(h->b_tables, tables.data(), std::max<size_t>(tables.size(), 1), s, "batch step tables");
    h->b_atoms.resize(std::max<size_t>(atoms_all.size(), 1));
    if (!atoms_all.empty()) copy_in(h->b_atoms, atoms_all.data(), atoms_all.size(), s, "batch atoms");
    h->b_fa32.resize(std::max<size_t>(fa_all.size(), 1));
    if (!fa_all.empty()) copy_in(h->b_fa32, fa_all.data(), fa_all.size(), s, "batch field arrays");
    h->b_mix_sigma.resize(std::max<size_t>(mix_sigma.size(), 1));
    if (!mix_sigma.empty()) copy_in(h->b_mix_sigma, mix_sigma.data(), mix_sigma.size(), s, "batch mixture cross sections");
    h->b_counters.resize(4ull * n_clouds);
    h->b_done.resize(n_clouds);
    h->b_start.resize(1);
    h->b_queue.resize(kBatchClasses);
    std::vector<IonPtrs> ions(n_clouds);
    for (uint32_t c = 0; c < n_clouds; ++c) {
        const size_t o = bc[c].off;
        IonPtrs& S = ions[c];
        S.x = h->b_x.p + o; S.y = h->b_y.p + o; S.z = h->b_z.p + o;
        S.vx = h->b_vx.p + o; S.vy = h->b_vy.p + o; S.vz = h->b_vz.p + o;
        S.mass = h->b_mass.p + o; S.charge = h->b_charge.p + o; S.species = h->b_spec.p + o;
        S.flags = h->b_flags.p + o;
        S.birth = bc[c].has_birth ? h->b_birth.p + o : nullptr;
        S.death = h->b_death.p + o; S.arrival = h->b_arrival.p + o; S.ncoll = h->b_ncoll.p + o;
        StepParams& P = params[c];
        P.species = h->b_species.p + species_off[c];
        P.table = h->b_tables.p + table_off[c];
        P.mix_sigma = mix_off[c] == SIZE_MAX ? nullptr : h->b_mix_sigma.p + mix_off[c];
        P.counters = h->b_counters.p + 4ull * c;
        P.atoms = h->b_atoms.p + atom_off[c];
        P.fa_f32 = fa_off[c] == SIZE_MAX ? nullptr : h->b_fa32.p + fa_off[c];
        P.fa_f64 = nullptr;   // FP64 clouds read the FP32 nodes, as after a dtype 0 upload
        P.mlp_blob = nullptr;
    }
    copy_in(h->b_params_d, params.data(), n_clouds, s, "batch params");
    copy_in(h->b_ions_d, ions.data(), n_clouds, s, "batch ion pointers");
    copy_in(h->b_kinds_d, kinds.data(), n_clouds, s, "batch kinds");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    h->b_clouds = std::move(bc);
    h->b_params = std::move(params);
    h->b_n = n_clouds;
    h->b_total = total;
    for (int k = 0; k < kBatchClasses; ++k) h->b_tiled_block[k] = 0;
}

// Tiles of one precision class, longest clouds first (step count x ions per
// step; FP64 bodies are costed like FP32 ones - the classes never share a
// queue).
void batch_build_tiles(icarion_gpu* h, int cls, uint32_t block) {
    if (h->b_tiled_block[cls] == block) return;
    std::vector<uint32_t> order;
    for (uint32_t c = 0; c < h->b_n; ++c)
        if (h->b_clouds[c].cls == cls && h->b_clouds[c].n)
            order.push_back(c);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return h->b_clouds[a].n_steps > h->b_clouds[b].n_steps;
    });
    std::vector<BatchTile> tiles;
    for (uint32_t c : order)
        for (size_t f = 0; f < h->b_clouds[c].n; f += block)
            tiles.push_back(BatchTile{c, static_cast<uint32_t>(f)});
    h->b_n_tiles[cls] = static_cast<uint32_t>(tiles.size());
    if (!tiles.empty()) {
        copy_in(h->b_tiles[cls], tiles.data(), tiles.size(), h->stream, "batch tiles");
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    }
    h->b_tiled_block[cls] = block;
}

icarion_batch_tuning batch_shape(const icarion_gpu* h, int cls, const icarion_batch_tuning* req) {
    icarion_batch_tuning t = req ? *req : h->b_tuned[cls];
    if (t.block_size == 0) t.block_size = 256;
    if (t.groups_per_cu == 0) t.groups_per_cu = 4;
    if (t.block_size != 64 && t.block_size != 128 && t.block_size != 256)
        fail(ICARION_ERR_INVALID, "batch block_size must be 64, 128 or 256");
    if (t.groups_per_cu > 64) fail(ICARION_ERR_INVALID, "batch groups_per_cu must be <= 64");
    return t;
}

// One batch launch captured as a HIP graph (reset + persistent kernels of the
// classes in `mask`); returns device milliseconds.
float batch_launch(icarion_gpu* h, icarion_batch_tuning t64, icarion_batch_tuning t32, int mask) {
    // the surrogate class has one shape: 256 ions per tile and two workgroups
    // per CU, the occupancy its registers (2 waves per SIMD) and 47 KB of LDS
    // weights allow
    const icarion_batch_tuning shape[kBatchClasses] = {t64, t32, {256, 2}};
    for (int cls = 0; cls < kBatchClasses; ++cls) batch_build_tiles(h, cls, shape[cls].block_size);
    const bool need_blob = ((mask >> 2) & 1) && h->b_n_tiles[2];
    if (need_blob && !h->have_surrogate)
        fail(ICARION_ERR_STATE, "the batch holds EHSS surrogate clouds but no surrogate parameters are set");
    BatchParams B[kBatchClasses];
    uint32_t block[kBatchClasses], grid[kBatchClasses];
    const uint32_t n_cu = static_cast<uint32_t>(std::max(h->n_cu, 1));
    for (int cls = 0; cls < kBatchClasses; ++cls) {
        block[cls] = shape[cls].block_size;
        grid[cls] = shape[cls].groups_per_cu * n_cu;
        B[cls].params = h->b_params_d.p;
        B[cls].ions = h->b_ions_d.p;
        B[cls].kinds = h->b_kinds_d.p;
        B[cls].tiles = h->b_tiles[cls].p;
        B[cls].n_tiles = (mask >> cls) & 1 ? h->b_n_tiles[cls] : 0;
        B[cls].with_mix = 0;
        for (const auto& c : h->b_clouds)
            if (c.cls == cls && (c.kind >= kBatchFirstMixKind || c.substeps)) B[cls].with_mix = 1;
        B[cls].queue = h->b_queue.p + cls;
        B[cls].done = h->b_done.p;
        B[cls].start = h->b_start.p;
    }
    hipStream_t s = h->stream;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    const hipError_t le = launch_batch(B, block, grid, h->mlp_blob.p, h->b_n, h->b_counters.p, s, h->side,
                                       h->ev_fork, h->ev_join);
    const hipError_t ee = hipStreamEndCapture(s, &graph);
    if (le != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        hip_check(le, "capture batch launch");
        hip_check(ee, "hipStreamEndCapture");
    }
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(graph);
        hip_check(ie, "hipGraphInstantiate");
    }
    hipError_t e = hipEventRecord(h->ev0, s);
    if (e == hipSuccess) e = hipGraphLaunch(exec, s);
    if (e == hipSuccess) e = hipEventRecord(h->ev1, s);
    if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev0, h->ev1);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    hip_check(e, "batch graph launch");
    return ms;
}

void do_batch_run(icarion_gpu* h, const icarion_batch_tuning* r64, const icarion_batch_tuning* r32,
                  icarion_cloud_stats* per_cloud, double* device_ms) {
    if (h->b_n == 0) fail(ICARION_ERR_STATE, "icarion_gpu_batch_upload not called");
    set_device(h);
    const icarion_batch_tuning t64 = batch_shape(h, 0, r64), t32 = batch_shape(h, 1, r32);
    const float ms = batch_launch(h, t64, t32, (1 << kBatchClasses) - 1);
    if (device_ms) *device_ms = ms;
    if (!per_cloud) return;
    const uint32_t n = h->b_n;
    std::vector<unsigned long long> cnt(4ull * n), done(n), start(1);
    std::vector<uint8_t> fl(h->b_total);
    std::vector<double> ar(h->b_total);
    hipStream_t s = h->stream;
    copy_out(cnt.data(), h->b_counters, cnt.size(), s, "batch counters");
    copy_out(done.data(), h->b_done, n, s, "batch completion clocks");
    copy_out(start.data(), h->b_start, 1, s, "batch start clock");
    copy_out(fl.data(), h->b_flags, h->b_total, s, "batch flags");
    copy_out(ar.data(), h->b_arrival, h->b_total, s, "batch arrival");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    const double tick_ms = h->clock_khz > 0 ? 1.0 / static_cast<double>(h->clock_khz) : 0.0;
    for (uint32_t c = 0; c < n; ++c) {
        const auto& bc = h->b_clouds[c];
        icarion_cloud_stats& st = per_cloud[c];
        st.ion_steps = cnt[4ull * c];
        st.collisions = cnt[4ull * c + 1];
        uint64_t alive = 0, arrived = 0;
        for (size_t i = bc.off; i < bc.off + bc.n; ++i) {
            if ((fl[i] & 3) == 3) ++alive;
            if (!std::isnan(ar[i])) ++arrived;
        }
        st.alive = alive;
        st.arrived = arrived;
        st.t_end_s = bc.t_end;
        st.done_ms = (done[c] && start[0] != ~0ull && done[c] >= start[0])
                         ? static_cast<double>(done[c] - start[0]) * tick_ms : 0.0;
    }
}

void do_batch_autotune(icarion_gpu* h, uint32_t trial_steps, icarion_batch_tuning* best64,
                       icarion_batch_tuning* best32, double* rate64, double* rate32) {
    if (h->b_n == 0) fail(ICARION_ERR_STATE, "icarion_gpu_batch_upload not called");
    if (trial_steps == 0) fail(ICARION_ERR_INVALID, "trial_steps must be > 0");
    set_device(h);
    hipStream_t s = h->stream;
    const size_t n = h->b_total;
    // snapshot of the mutable state, restored before every trial and at the end
    DevBuf<double> sx, sy, sz, svx, svy, svz, sd, sa;
    DevBuf<uint32_t> snc;
    DevBuf<uint8_t> sfl;
    auto d2d = [&](auto& dst, const auto& src) {
        if (n) hip_check(hipMemcpyAsync(dst.p, src.p, n * sizeof(*src.p), hipMemcpyDeviceToDevice, s), "batch snapshot");
    };
    sx.resize(n); sy.resize(n); sz.resize(n); svx.resize(n); svy.resize(n); svz.resize(n); sd.resize(n);
    sa.resize(n); snc.resize(n); sfl.resize(n);
    d2d(sx, h->b_x); d2d(sy, h->b_y); d2d(sz, h->b_z); d2d(svx, h->b_vx); d2d(svy, h->b_vy); d2d(svz, h->b_vz);
    d2d(sd, h->b_death); d2d(sa, h->b_arrival); d2d(snc, h->b_ncoll); d2d(sfl, h->b_flags);
    auto restore = [&]() {
        d2d(h->b_x, sx); d2d(h->b_y, sy); d2d(h->b_z, sz); d2d(h->b_vx, svx); d2d(h->b_vy, svy);
        d2d(h->b_vz, svz); d2d(h->b_death, sd); d2d(h->b_arrival, sa); d2d(h->b_ncoll, snc);
        d2d(h->b_flags, sfl);
    };
    // trial parameters: every cloud cut to at most trial_steps steps
    std::vector<StepParams> trial = h->b_params;
    for (auto& P : trial) P.n_steps = std::min<uint32_t>(P.n_steps, trial_steps);
    copy_in(h->b_params_d, trial.data(), trial.size(), s, "batch trial params");
    static const uint32_t kBlocks[] = {64, 128, 256};
    static const uint32_t kGroups[] = {1, 2, 4, 8, 16};
    icarion_batch_tuning best[2] = {{256, 4}, {256, 4}};
    double best_rate[2] = {0.0, 0.0};
    for (int cls = 0; cls < 2; ++cls) {
        bool present = false;
        for (const auto& c : h->b_clouds) present |= c.cls == cls;
        if (!present) continue;
        for (uint32_t b : kBlocks) {
            for (uint32_t g : kGroups) {
                if (static_cast<uint64_t>(b) * g > 2048) continue;   // resident-thread limit per CU
                restore();
                const icarion_batch_tuning t{b, g};
                const float ms = batch_launch(h, t, t, 1 << cls);
                std::vector<unsigned long long> cnt(4ull * h->b_n);
                copy_out(cnt.data(), h->b_counters, cnt.size(), s, "batch counters");
                hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
                uint64_t steps = 0;
                for (uint32_t c = 0; c < h->b_n; ++c) steps += cnt[4ull * c];
                const double rate = ms > 0.0f ? static_cast<double>(steps) / (ms * 1e-3) : 0.0;
                if (rate > best_rate[cls]) {
                    best_rate[cls] = rate;
                    best[cls] = t;
                }
            }
        }
        h->b_tuned[cls] = best[cls];
    }
    restore();
    copy_in(h->b_params_d, h->b_params.data(), h->b_params.size(), s, "batch params");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (best64) *best64 = best[0];
    if (best32) *best32 = best[1];
    if (rate64) *rate64 = best_rate[0];
    if (rate32) *rate32 = best_rate[1];
}

void do_batch_download(icarion_gpu* h, const icarion_cloud_desc* clouds, uint32_t n_clouds) {
    if (h->b_n == 0) fail(ICARION_ERR_STATE, "no resident batch");
    if (!clouds || n_clouds != h->b_n) fail(ICARION_ERR_INVALID, "n_clouds differs from the resident batch");
    set_device(h);
    hipStream_t s = h->stream;
    const size_t n = h->b_total;
    std::vector<double> hx(n), hy(n), hz(n), hvx(n), hvy(n), hvz(n), hd(n), ha(n);
    std::vector<uint32_t> hn(n);
    std::vector<uint8_t> hf(n);
    copy_out(hx.data(), h->b_x, n, s, "batch x"); copy_out(hy.data(), h->b_y, n, s, "batch y");
    copy_out(hz.data(), h->b_z, n, s, "batch z"); copy_out(hvx.data(), h->b_vx, n, s, "batch vx");
    copy_out(hvy.data(), h->b_vy, n, s, "batch vy"); copy_out(hvz.data(), h->b_vz, n, s, "batch vz");
    copy_out(hd.data(), h->b_death, n, s, "batch death"); copy_out(ha.data(), h->b_arrival, n, s, "batch arrival");
    copy_out(hn.data(), h->b_ncoll, n, s, "batch ncoll"); copy_out(hf.data(), h->b_flags, n, s, "batch flags");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    for (uint32_t c = 0; c < n_clouds; ++c) {
        icarion_ion_soa* io = clouds[c].ions;
        const auto& bc = h->b_clouds[c];
        if (!io || io->n != bc.n) fail(ICARION_ERR_INVALID, "cloud " + std::to_string(c) + ": ions->n differs");
        const size_t o = bc.off;
        for (size_t i = 0; i < bc.n; ++i) {
            if (io->x) io->x[i] = hx[o + i];
            if (io->y) io->y[i] = hy[o + i];
            if (io->z) io->z[i] = hz[o + i];
            if (io->vx) io->vx[i] = hvx[o + i];
            if (io->vy) io->vy[i] = hvy[o + i];
            if (io->vz) io->vz[i] = hvz[o + i];
            if (io->death_time_s) io->death_time_s[i] = hd[o + i];
            if (io->arrival_time_s) io->arrival_time_s[i] = ha[o + i];
            if (io->n_collisions) io->n_collisions[i] = hn[o + i];
            if (io->active) io->active[i] = hf[o + i] & 1;
            if (io->born) io->born[i] = (hf[o + i] >> 1) & 1;
        }
    }
}

}  // namespace

namespace {

// The 64 orientations of EHSSCollisionHandler::compute_effective_ccs: PhysicsRng(12345)
// (std::mt19937_64 + uniform_real_distribution) through generate_random_rotation
// (CollisionGeometry.cpp:30-59).
std::vector<double> oapa_rotations() {
    std::mt19937_64 eng(12345);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    std::vector<double> R(64 * 9);
    for (int k = 0; k < 64; ++k) {
        const double u1 = uni(eng), u2 = uni(eng), u3 = uni(eng);
        const double x = std::sqrt(1.0 - u1) * std::sin(2.0 * kPiHost * u2);
        const double y = std::sqrt(1.0 - u1) * std::cos(2.0 * kPiHost * u2);
        const double z = std::sqrt(u1) * std::sin(2.0 * kPiHost * u3);
        const double w = std::sqrt(u1) * std::cos(2.0 * kPiHost * u3);
        double* m = &R[9 * k];
        m[0] = 1 - 2 * (y * y + z * z); m[1] = 2 * (x * y - z * w);     m[2] = 2 * (x * z + y * w);
        m[3] = 2 * (x * y + z * w);     m[4] = 1 - 2 * (x * x + z * z); m[5] = 2 * (y * z - x * w);
        m[6] = 2 * (x * z - y * w);     m[7] = 2 * (y * z + x * w);     m[8] = 1 - 2 * (x * x + y * y);
    }
    return R;
}

void install_params(icarion_gpu* h, const float* params) {
    h->params.assign(params, params + kMlpParams);
    std::vector<unsigned char> blob(kPackBytes);
    mlp_pack_blob(h->params.data(), blob.data());
    set_device(h);
    copy_in(h->mlp_blob, reinterpret_cast<const uint4*>(blob.data()), kPackBytes / 16, h->stream, "upload surrogate");
    hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    h->have_surrogate = true;
}

std::vector<float> init_params(uint64_t seed) {
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases
    std::mt19937_64 eng(seed);
    std::vector<float> p(kMlpParams, 0.f);
    auto fill = [&](int off, int n, int fan_in) {
        std::uniform_real_distribution<float> u(-1.f / std::sqrt((float)fan_in), 1.f / std::sqrt((float)fan_in));
        for (int i = 0; i < n; ++i) p[off + i] = u(eng);
    };
    fill(kOffW1, kMlpH * kMlpIn, kMlpIn);
    fill(kOffW2, kMlpH * kMlpH, kMlpH);
    fill(kOffW3, kMlpOut * kMlpH, kMlpH);
    return p;
}

__global__ void gather_batch_kernel(const float* X, const float* T, const uint32_t* perm, const uint32_t* ctr,
                                    float* Xb, float* Tb, int B) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= B) return;
    const size_t src = perm[static_cast<size_t>(ctr[0]) * B + r];
    for (int j = 0; j < kMlpIn; ++j) Xb[r * kMlpIn + j] = X[src * kMlpIn + j];
    for (int j = 0; j < kMlpOut; ++j) Tb[r * kMlpOut + j] = T[src * kMlpOut + j];
}

__global__ void adam_graph_kernel(float* p, const float* g, float* m, float* v, int n, float lr, float b1,
                                  float b2, float eps, const uint32_t* ctr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float step = static_cast<float>(ctr[1] + 1);
    const float c1 = 1.f / (1.f - powf(b1, step));
    const float c2 = 1.f / (1.f - powf(b2, step));
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void advance_counters_kernel(uint32_t* ctr) {
    if (threadIdx.x == 0) {
        ctr[0] += 1;
        ctr[1] += 1;
    }
}

__global__ void zero_f32_kernel(float* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ void zero_u32_kernel(uint32_t* p, int n) {
    const int i = threadIdx.x;
    if (i < n) p[i] = 0u;
}

struct TrainDev {
    DevBuf<float> Z1, H1, Z2, H2, Y, dY, dH2, dH1, params, grad, m, v, loss, Xb, Tb, ws;
    void alloc(int B) {
        Z1.resize(static_cast<size_t>(B) * kMlpH); H1.resize(static_cast<size_t>(B) * kMlpH);
        Z2.resize(static_cast<size_t>(B) * kMlpH); H2.resize(static_cast<size_t>(B) * kMlpH);
        Y.resize(static_cast<size_t>(B) * kMlpOut); dY.resize(static_cast<size_t>(B) * kMlpOut);
        dH2.resize(static_cast<size_t>(B) * kMlpH); dH1.resize(static_cast<size_t>(B) * kMlpH);
        params.resize(kMlpParams); grad.resize(kMlpParams); m.resize(kMlpParams); v.resize(kMlpParams);
        loss.resize(1);
        Xb.resize(static_cast<size_t>(B) * kMlpIn); Tb.resize(static_cast<size_t>(B) * kMlpOut);
        ws.resize(kGemmWorkspaceFloats);
    }
    TrainBuffers view(float* X, float* T, int B) {
        TrainBuffers t{};
        t.X = X; t.T = T; t.Z1 = Z1.p; t.H1 = H1.p; t.Z2 = Z2.p; t.H2 = H2.p; t.Y = Y.p; t.dY = dY.p;
        t.dH2 = dH2.p; t.dH1 = dH1.p; t.params = params.p; t.grad = grad.p; t.m = m.p; t.v = v.p;
        t.loss = loss.p; t.ws = ws.p; t.B = B;
        return t;
    }
    void release() {
        for (DevBuf<float>* b : {&Z1, &H1, &Z2, &H2, &Y, &dY, &dH2, &dH1, &params, &grad, &m, &v, &loss, &Xb, &Tb, &ws})
            b->release();
    }
};

}  // namespace

extern "C" {

int icarion_gpu_abi_version(void) { return ICARION_GPU_ABI_VERSION; }

int icarion_gpu_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

size_t icarion_gpu_last_error(char* buf, size_t len) {
    if (buf && len) {
        const size_t k = std::min(len - 1, g_last_error.size());
        std::memcpy(buf, g_last_error.data(), k);
        buf[k] = '\0';
    }
    return g_last_error.size();
}

int icarion_gpu_create(const icarion_gpu_desc* desc, icarion_gpu** out) {
    return guarded([&] {
        if (!out) fail(ICARION_ERR_INVALID, "out is NULL");
        *out = nullptr;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            fail(ICARION_ERR_NO_DEVICE, "no HIP device visible");
        auto h = new icarion_gpu();
        try {
            if (desc) {
                h->device = desc->device;
                if (desc->block_size) h->block = desc->block_size;
                h->steps_per_launch = desc->steps_per_launch;
            }
            if (h->device < 0 || h->device >= count) fail(ICARION_ERR_INVALID, "device ordinal out of range");
            if (h->block % 64 != 0 || h->block > 256 || h->block == 0)
                fail(ICARION_ERR_INVALID, "block_size must be 64, 128, 192 or 256");
            set_device(h);
            hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreate(&h->ev0), "hipEventCreate");
            hip_check(hipEventCreate(&h->ev1), "hipEventCreate");
            h->counters.resize(kCounterWords);
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&h->host_ctr), 4 * sizeof(unsigned long long),
                                    hipHostMallocMapped | hipHostMallocCoherent),
                      "hipHostMalloc counters");
            hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&h->host_ctr_dev), h->host_ctr, 0),
                      "hipHostGetDevicePointer");
            if (h->device < kMaxScratchDevices) {
                // once per device, with that device current; a failed lookup
                // is kept and reported by icarion_gpu_step_noalloc, the one
                // entry point that needs the address
                std::lock_guard<std::mutex> lock(g_scratch_mu);
                if (!g_scratch[h->device] && g_scratch_error[h->device].empty()) {
                    void* ptr = nullptr;
                    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
                    if (e != hipSuccess) {
                        g_scratch_error[h->device] =
                            std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
                        (void)hipGetLastError();
                    } else if (!ptr) {
                        g_scratch_error[h->device] = "hipGetSymbolAddress(scratch) returned NULL";
                    } else {
                        g_scratch[h->device] = ptr;
                    }
                }
                h->scratch_dev = static_cast<unsigned char*>(g_scratch[h->device]);
            }
            hip_check(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming), "hipEventCreate");
            for (int k = 0; k < kBatchClasses - 1; ++k) {
                hip_check(hipStreamCreateWithFlags(&h->side[k], hipStreamNonBlocking), "hipStreamCreate");
                hip_check(hipEventCreateWithFlags(&h->ev_join[k], hipEventDisableTiming), "hipEventCreate");
            }
            hip_check(hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, h->device),
                      "hipDeviceGetAttribute");
            hip_check(hipDeviceGetAttribute(&h->clock_khz, hipDeviceAttributeWallClockRate, h->device),
                      "hipDeviceGetAttribute");
        } catch (...) {
            icarion_gpu_destroy(h);
            throw;
        }
        *out = h;
    });
}

int icarion_gpu_destroy(icarion_gpu* h) {
    if (!h) return ICARION_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (hipStream_t st : h->side)
        if (st) (void)hipStreamSynchronize(st);
    h->species.release(); h->atoms_f32.release(); h->atom_offsets.release(); h->atom_counts.release();
    h->mlp_blob.release(); h->fa32.release(); h->fa64.release();
    for (auto& F : h->dom_field) {
        F.fa32.release();
        F.fa64.release();
    }
    h->dom_species_d.release();
    for (auto& a : h->ax64) a.release();
    for (auto& a : h->ax32) a.release();
    h->ord.release(); h->ord_keys_a.release(); h->ord_keys_b.release(); h->ord_idx.release(); h->ord_temp.release();
    h->mix_sigma.release(); h->emix_sigma.release();
    h->off_tab.release(); h->off_sigma.release(); h->off_mu.release();
    h->ori_tab.release(); h->ori_R.release(); h->ori_area.release();
    h->r_t.release(); h->r_dt.release(); h->r_err.release(); h->r_k1x.release(); h->r_k1y.release();
    h->r_k1z.release(); h->r_fsal.release(); h->r_eng.release();
    h->b_params_d.release(); h->b_ions_d.release(); h->b_kinds_d.release(); h->b_queue.release();
    for (auto& t : h->b_tiles) t.release();
    h->b_done.release(); h->b_start.release();
    h->b_counters.release(); h->b_x.release(); h->b_y.release(); h->b_z.release(); h->b_vx.release();
    h->b_vy.release(); h->b_vz.release(); h->b_mass.release(); h->b_charge.release(); h->b_birth.release();
    h->b_death.release(); h->b_arrival.release(); h->b_spec.release(); h->b_ncoll.release();
    h->b_flags.release(); h->b_species.release(); h->b_tables.release();
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    for (hipEvent_t ev : h->ev_join)
        if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t st : h->side)
        if (st) (void)hipStreamDestroy(st);
    h->x.release(); h->y.release(); h->z.release(); h->vx.release(); h->vy.release(); h->vz.release();
    h->mass.release(); h->charge.release(); h->birth.release(); h->death.release(); h->arrival.release();
    h->spec.release(); h->ncoll.release(); h->flags.release(); h->table.release(); h->counters.release();
    h->term_scales.release();
    h->dom_params_d.release(); h->dom_tables.release(); h->dom_spans_d.release(); h->dom_ap_d.release();
    h->dom_mix_d.release(); h->dom_emix_d.release();
    h->ion_dom.release();
    for (int j = 0; j < 3; ++j) {
        h->snap.full[j].release();
        h->snap.rec[j].release();
        if (h->snap.packed[j]) (void)hipEventDestroy(h->snap.packed[j]);
    }
    if (h->snap.copied) (void)hipEventDestroy(h->snap.copied);
    h->snap.count.release();
    h->snap.prev.release();
    if (h->snap.host_full) (void)hipHostFree(h->snap.host_full);
    if (h->snap.host_rec) (void)hipHostFree(h->snap.host_rec);
    if (h->snap.host_count) (void)hipHostFree(h->snap.host_count);
    if (h->host_ctr) (void)hipHostFree(h->host_ctr);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return ICARION_OK;
}

int icarion_gpu_set_domain(icarion_gpu* h, const icarion_domain_desc* domain,
                           const icarion_physics_desc* physics) {
    return guarded([&] {
        if (!h || !domain || !physics) fail(ICARION_ERR_INVALID, "NULL argument");
        validate_domain(*domain, *physics);
        h->dom = *domain;
        h->phys = *physics;
        h->base = derive_params(*domain, *physics);
        h->have_domain = true;
        h->n_dom = 0;
    });
}

int icarion_gpu_set_domains(icarion_gpu* h, const icarion_domain_desc* domains, const double* end_aperture_m,
                            uint32_t n_domains, const icarion_physics_desc* physics) {
    return guarded([&] {
        if (!h || !domains || !physics) fail(ICARION_ERR_INVALID, "NULL argument");
        if (n_domains == 0) fail(ICARION_ERR_INVALID, "n_domains must be >= 1");
        if (n_domains > ICARION_MAX_DOMAINS) fail(ICARION_ERR_UNSUPPORTED, "at most ICARION_MAX_DOMAINS domains");
        for (uint32_t k = 0; k < n_domains; ++k) {
            validate_domain(domains[k], *physics);
            if (end_aperture_m && !(end_aperture_m[k] >= 0.0))
                fail(ICARION_ERR_INVALID, "end_aperture_m must be >= 0");
        }
        h->dom = domains[0];
        h->phys = *physics;
        h->base = derive_params(domains[0], *physics);
        h->have_domain = true;
        h->n_dom = 0;
        h->doms.clear();
        h->dom_ap_sq.clear();
        h->dom_spans.clear();
        h->dom_has_orbitrap = false;
        h->dom_wave.clear();
        h->dom_pressure.clear();
        h->dom_flow.clear();
        h->dom_mix.clear();
        h->dom_mix_sig.clear();
        h->dom_mix_ns.clear();
        h->dom_mix_all.clear();
        h->dom_emix_sig.clear();
        h->dom_species.clear();
        for (auto& F : h->dom_field) {   // DevBuf frees only on release()
            F.fa32.release();
            F.fa64.release();
        }
        h->dom_field.clear();
        if (n_domains == 1) return;
        h->doms.assign(domains, domains + n_domains);
        h->dom_wave.resize(n_domains);
        h->dom_pressure.resize(n_domains);
        h->dom_flow.assign(n_domains, {ICARION_FLOW_CONSTANT, 0.0});
        h->dom_mix.resize(n_domains);
        h->dom_mix_sig.resize(n_domains);
        h->dom_mix_ns.assign(n_domains, 0);
        h->dom_mix_all.resize(n_domains);
        h->dom_emix_sig.resize(n_domains);
        h->dom_species.resize(n_domains);
        h->dom_field.resize(n_domains);
        // DomainManager's constructor (DomainManager.cpp:26-63): a span per
        // cylindrical domain, z range and radius widened by 1e-12 m, with the
        // index of its domain, sorted by z_min (std::sort, as there)
        constexpr double kDomainEps = 1e-12;   // DOMAIN_BOUNDARY_EPSILON (DomainManager.h:42)
        for (uint32_t k = 0; k < n_domains; ++k) {
            const icarion_domain_desc& d = domains[k];
            const double ap = (end_aperture_m && end_aperture_m[k] > 0.0) ? end_aperture_m[k] : d.radius_m;
            h->dom_ap_sq.push_back(ap * ap);
            if (d.instrument == ICARION_INSTRUMENT_ORBITRAP) {
                h->dom_has_orbitrap = true;
                continue;
            }
            const double z0 = d.origin_m[2], z1 = z0 + d.length_m;
            DomainSpan sp{};
            sp.z_min = std::min(z0, z1) - kDomainEps;
            sp.z_max = std::max(z0, z1) + kDomainEps;
            sp.radius = d.radius_m + kDomainEps;
            sp.idx = static_cast<int>(k);
            h->dom_spans.push_back(sp);
        }
        std::sort(h->dom_spans.begin(), h->dom_spans.end(),
                  [](const DomainSpan& a, const DomainSpan& b) { return a.z_min < b.z_min; });
        h->n_dom = n_domains;
    });
}

int icarion_gpu_arrival_stats(icarion_gpu* h, icarion_arrival_stats* stats) {
    return guarded([&] {
        if (!h || !stats) fail(ICARION_ERR_INVALID, "NULL argument");
        if (!h->have_state) fail(ICARION_ERR_STATE, "no resident state");
        set_device(h);
        // per workgroup {sum, sum of squares, min, max, count}, summed here in order
        struct Partial {
            double sum, sumsq, min, max;
            unsigned long long count, pad;
        };
        const uint32_t n = static_cast<uint32_t>(h->n);
        const uint32_t nb = arrival_blocks(n);
        DevBuf<Partial> part;
        std::vector<Partial> hp(nb);
        if (nb) {
            part.resize(nb);
            hip_check(launch_arrival_partials(h->arrival.p, n, part.p, h->stream), "arrival statistics launch");
            copy_out(hp.data(), part, nb, h->stream, "read arrival statistics");
            hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        }
        double sum = 0.0, sq = 0.0, mn = INFINITY, mx = -INFINITY;
        uint64_t cnt = 0;
        for (const Partial& p : hp) {
            sum += p.sum;
            sq += p.sumsq;
            mn = std::min(mn, p.min);
            mx = std::max(mx, p.max);
            cnt += p.count;
        }
        stats->count = cnt;
        stats->min_s = cnt ? mn : NAN;
        stats->max_s = cnt ? mx : NAN;
        stats->mean_s = cnt ? sum / static_cast<double>(cnt) : NAN;
        const double var = cnt ? sq / static_cast<double>(cnt) - stats->mean_s * stats->mean_s : NAN;
        stats->std_s = cnt ? std::sqrt(std::max(0.0, var)) : NAN;
    });
}

int icarion_gpu_arrival_histogram(icarion_gpu* h, double lo_s, double hi_s, uint32_t n_bins, uint64_t* counts) {
    return guarded([&] {
        if (!h || !counts) fail(ICARION_ERR_INVALID, "NULL argument");
        if (!h->have_state) fail(ICARION_ERR_STATE, "no resident state");
        if (n_bins == 0 || n_bins > ICARION_MAX_HIST_BINS)
            fail(ICARION_ERR_INVALID, "n_bins must be in 1 .. ICARION_MAX_HIST_BINS");
        if (!(lo_s < hi_s) || !std::isfinite(lo_s) || !std::isfinite(hi_s))
            fail(ICARION_ERR_INVALID, "need finite lo_s < hi_s");
        set_device(h);
        DevBuf<unsigned long long> bins;
        bins.resize(n_bins);
        hip_check(launch_arrival_histogram(h->arrival.p, static_cast<uint32_t>(h->n), lo_s, hi_s, n_bins, bins.p,
                                           h->stream),
                  "arrival histogram launch");
        copy_out(reinterpret_cast<unsigned long long*>(counts), bins, n_bins, h->stream, "read arrival histogram");
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    });
}

int icarion_gpu_download_domains(icarion_gpu* h, int32_t* domain_index) {
    return guarded([&] {
        if (!h || (!domain_index && h->n)) fail(ICARION_ERR_INVALID, "NULL argument");
        if (!h->have_state) fail(ICARION_ERR_STATE, "no resident state");
        set_device(h);
        if (h->n) copy_out(domain_index, h->ion_dom, h->n, h->stream, "download domain index");
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    });
}

int icarion_gpu_set_rk45(icarion_gpu* h, const icarion_rk45_desc* d) {
    return guarded([&] {
        if (!h || !d) fail(ICARION_ERR_INVALID, "NULL argument");
        // RK45Strategy(const AdaptiveConfig&) (RK45Strategy.cpp:179-195)
        if (!(d->atol > 0.0) || !(d->rtol > 0.0)) fail(ICARION_ERR_INVALID, "RK45 tolerances must be positive");
        if (!(d->safety_factor > 0.0 && d->safety_factor < 1.0))
            fail(ICARION_ERR_INVALID, "RK45 safety factor must be in (0, 1)");
        if (!(d->absolute_min_step_s >= 0.0))
            fail(ICARION_ERR_INVALID, "RK45 absolute_min_step_s must be non-negative");
        if (!(d->min_step_factor > 0.0 && d->max_step_factor > 0.0 && d->max_step_increase > 0.0 &&
              d->max_step_decrease > 0.0))
            fail(ICARION_ERR_INVALID, "RK45 step factors must be positive");
        h->rk45 = Rk45Cfg{d->atol, d->rtol, d->safety_factor, d->min_step_factor, d->max_step_factor,
                          d->max_step_increase, d->max_step_decrease, d->absolute_min_step_s,
                          d->accept_at_dt_min ? 1 : 0, 0};
    });
}

namespace {

// Device rows of a species table; every atom range must lie in the n_atoms
// atoms of the handle's table.
std::vector<SpeciesDev> species_rows(const icarion_species_desc* species, uint32_t n_species, uint64_t n_atoms) {
    std::vector<SpeciesDev> sd(n_species);
    for (uint32_t s = 0; s < n_species; ++s) {
        const icarion_species_desc& d = species[s];
        if (static_cast<uint64_t>(d.atom_offset) + d.atom_count > n_atoms)
            fail(ICARION_ERR_INVALID, "species " + std::to_string(s) + " atoms exceed the atom table");
        sd[s].sigma_hss = d.sigma_hss_m2;
        sd[s].sigma_ehss = d.sigma_ehss_m2;
        sd[s].K0_cm2_Vs = d.K0_cm2_Vs;
        sd[s].atom_offset = d.atom_offset;
        sd[s].atom_count = d.atom_count;
        for (int f = 0; f < 3; ++f) sd[s].feat[f] = d.surrogate_features[f];
        sd[s].pad = 0.f;
    }
    return sd;
}

}  // namespace

int icarion_gpu_set_species(icarion_gpu* h, const icarion_species_desc* species, uint32_t n_species,
                            const void* atoms_xyzr, uint32_t n_atoms, int32_t atoms_dtype) {
    return guarded([&] {
        if (!h || !species || n_species == 0) fail(ICARION_ERR_INVALID, "species table is empty");
        if (n_atoms && !atoms_xyzr) fail(ICARION_ERR_INVALID, "atoms_xyzr is NULL");
        if (atoms_dtype != 0 && atoms_dtype != 1) fail(ICARION_ERR_INVALID, "atoms_dtype must be 0 or 1");
        // The device holds the atom table in float only; a double table is
        // declined rather than rounded behind the caller's back.
        if (n_atoms && atoms_dtype == 1)
            fail(ICARION_ERR_UNSUPPORTED, "the device atom table is float (atoms_dtype 0); round the "
                                          "coordinates to float or use the CPU path for double atoms");
        const std::vector<SpeciesDev> sd = species_rows(species, n_species, n_atoms);
        std::vector<uint32_t> offs(n_species), cnts(n_species);
        for (uint32_t s = 0; s < n_species; ++s) {
            offs[s] = species[s].atom_offset;
            cnts[s] = species[s].atom_count;
        }
        set_device(h);
        copy_in(h->species, sd.data(), n_species, h->stream, "upload species");
        copy_in(h->atom_offsets, offs.data(), n_species, h->stream, "upload atom offsets");
        copy_in(h->atom_counts, cnts.data(), n_species, h->stream, "upload atom counts");
        h->atoms_f32.release();
        if (n_atoms) {
            copy_in(h->atoms_f32, static_cast<const float4*>(atoms_xyzr), n_atoms, h->stream, "upload atoms");
        }
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        h->species_host.assign(species, species + n_species);
        h->n_species = n_species;
        h->n_atoms = n_atoms;
        h->have_species = true;
    });
}

namespace {

// Node-major device copies of the terms (element node * n_terms + term);
// double copies too when any term arrives in double.
void upload_field_data_to(icarion_gpu* h, DevBuf<float4>& fa32, DevBuf<double4>& fa64, size_t nn,
                          const icarion_field_term_desc* terms, uint32_t nt) {
    bool any_double = false;
    for (uint32_t j = 0; j < nt; ++j) any_double |= terms[j].dtype == 1;
    std::vector<float4> f(nn * nt);
    std::vector<double4> g(any_double ? nn * nt : 0);
    for (uint32_t j = 0; j < nt; ++j) {
        if (terms[j].dtype == 0) {
            const float* e = static_cast<const float*>(terms[j].E);
            for (size_t i = 0; i < nn; ++i) {
                f[i * nt + j] = make_float4(e[3 * i], e[3 * i + 1], e[3 * i + 2], 0.0f);
                if (any_double) g[i * nt + j] = make_double4(e[3 * i], e[3 * i + 1], e[3 * i + 2], 0.0);
            }
        } else {
            const double* e = static_cast<const double*>(terms[j].E);
            for (size_t i = 0; i < nn; ++i) {
                f[i * nt + j] = make_float4((float)e[3 * i], (float)e[3 * i + 1], (float)e[3 * i + 2], 0.0f);
                g[i * nt + j] = make_double4(e[3 * i], e[3 * i + 1], e[3 * i + 2], 0.0);
            }
        }
    }
    if (any_double) copy_in(fa64, g.data(), g.size(), h->stream, "upload field (double)");
    else fa64.release();
    copy_in(fa32, f.data(), f.size(), h->stream, "upload field (float)");
    hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
}

void upload_field_data(icarion_gpu* h, size_t nn, const icarion_field_term_desc* terms, uint32_t nt) {
    upload_field_data_to(h, h->fa32, h->fa64, nn, terms, nt);
}

// An axis whose nodes are x0 + i * d to within 1e-12 of its span.
bool uniform_axis(const double* a, uint32_t n, double& d) {
    d = (a[n - 1] - a[0]) / static_cast<double>(n - 1);
    if (!(d > 0.0)) return false;
    const double tol = 1e-12 * (a[n - 1] - a[0]);
    for (uint32_t i = 0; i < n; ++i)
        if (std::fabs(a[i] - (a[0] + static_cast<double>(i) * d)) > tol) return false;
    return true;
}

}  // namespace

int icarion_gpu_upload_field_array(icarion_gpu* h, uint32_t nx, uint32_t ny, uint32_t nz, double x0,
                                   double dx, double y0, double dy, double z0, double dz,
                                   const void* E, int32_t dtype) {
    return guarded([&] {
        if (!h || !E) fail(ICARION_ERR_INVALID, "NULL argument");
        if (nx < 2 || ny < 2 || nz < 2) fail(ICARION_ERR_INVALID, "field array needs >= 2 nodes per axis");
        if (!(dx > 0.0 && dy > 0.0 && dz > 0.0)) fail(ICARION_ERR_INVALID, "grid spacing must be > 0");
        if (dtype != 0 && dtype != 1) fail(ICARION_ERR_INVALID, "dtype must be 0 (float) or 1 (double)");
        set_device(h);
        icarion_field_term_desc t{};
        t.E = E;
        t.dtype = dtype;
        upload_field_data(h, static_cast<size_t>(nx) * ny * nz, &t, 1);
        for (auto& a : h->ax64) a.release();
        for (auto& a : h->ax32) a.release();
        icarion_gpu::FieldSrc f;
        f.nx = static_cast<int>(nx); f.ny = static_cast<int>(ny); f.nz = static_cast<int>(nz);
        f.n_terms = 1;
        f.x0 = x0; f.y0 = y0; f.z0 = z0;
        f.dx = dx; f.dy = dy; f.dz = dz;
        f.x_last = x0 + (nx - 1) * dx;
        f.y_last = y0 + (ny - 1) * dy;
        f.z_last = z0 + (nz - 1) * dz;
        h->fsrc = f;
        h->have_field = true;
    });
}

int icarion_gpu_upload_domain_field_array(icarion_gpu* h, uint32_t domain, uint32_t nx, uint32_t ny, uint32_t nz,
                                          double x0, double dx, double y0, double dy, double z0, double dz,
                                          const void* E, int32_t dtype) {
    return guarded([&] {
        if (!h || !E) fail(ICARION_ERR_INVALID, "NULL argument");
        if (h->n_dom < 2) fail(ICARION_ERR_STATE, "icarion_gpu_set_domains with two or more domains not called");
        if (domain >= h->n_dom) fail(ICARION_ERR_INVALID, "domain index out of range");
        if (!h->doms[domain].use_field_array)
            fail(ICARION_ERR_INVALID, "the domain's use_field_array is 0");
        if (nx < 2 || ny < 2 || nz < 2) fail(ICARION_ERR_INVALID, "field array needs >= 2 nodes per axis");
        if (!(dx > 0.0 && dy > 0.0 && dz > 0.0)) fail(ICARION_ERR_INVALID, "grid spacing must be > 0");
        if (dtype != 0 && dtype != 1) fail(ICARION_ERR_INVALID, "dtype must be 0 (float) or 1 (double)");
        set_device(h);
        icarion_field_term_desc t{};
        t.E = E;
        t.dtype = dtype;
        auto& F = h->dom_field[domain];
        upload_field_data_to(h, F.fa32, F.fa64, static_cast<size_t>(nx) * ny * nz, &t, 1);
        F.nx = static_cast<int>(nx); F.ny = static_cast<int>(ny); F.nz = static_cast<int>(nz);
        F.x0 = x0; F.y0 = y0; F.z0 = z0;
        F.dx = dx; F.dy = dy; F.dz = dz;
        F.have = true;
    });
}

int icarion_gpu_set_gas_mixture(icarion_gpu* h, const icarion_gas_component* components,
                                uint32_t n_components, const double* sigma_m2, uint32_t n_species) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        if (n_components == 0) {
            h->mix.clear();
            h->mix_all.clear();
            h->mix_radius.clear();
            h->mix_species = 0;
            h->mix_sigma.release();
            h->emix_sigma.release();
            return;
        }
        if (!components || !sigma_m2) fail(ICARION_ERR_INVALID, "NULL argument");
        if (!h->have_species || n_species != h->n_species)
            fail(ICARION_ERR_INVALID, "n_species must match the species table (set_species first)");
        std::vector<icarion_gas_component> mix;
        std::vector<double> sig;
        build_hss_mixture(components, n_components, sigma_m2, n_species, mix, sig);
        set_device(h);
        copy_in(h->mix_sigma, sig.data(), sig.size(), h->stream, "upload mixture cross sections");
        copy_in(h->emix_sigma, sigma_m2, static_cast<size_t>(n_species) * n_components, h->stream,
                "upload mixture cross sections");
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        h->mix = std::move(mix);
        h->mix_all.assign(components, components + n_components);
        h->mix_radius.assign(n_components, 0.0);
        h->mix_species = n_species;
    });
}

int icarion_gpu_set_gas_radii(icarion_gpu* h, const double* radius_m, uint32_t n_components) {
    return guarded([&] {
        if (!h || !radius_m) fail(ICARION_ERR_INVALID, "NULL argument");
        if (n_components != h->mix_all.size())
            fail(ICARION_ERR_INVALID, "n_components must match the gas mixture (set_gas_mixture first)");
        for (uint32_t c = 0; c < n_components; ++c)
            if (!(radius_m[c] >= 0.0) || !std::isfinite(radius_m[c]))
                fail(ICARION_ERR_INVALID, "radii must be finite and >= 0");
        h->mix_radius.assign(radius_m, radius_m + n_components);
    });
}

int icarion_gpu_set_waveforms(icarion_gpu* h, const icarion_waveform_desc* const* slots) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        std::unique_ptr<HostWaveform> w[ICARION_N_VOLTAGE_SLOTS];
        const bool any = parse_waveform_slots(slots, w);
        for (int k = 0; k < ICARION_N_VOLTAGE_SLOTS; ++k) h->wave[k] = std::move(w[k]);
        h->has_waves = any;
    });
}

int icarion_gpu_set_domain_gas_mixture(icarion_gpu* h, uint32_t domain, const icarion_gas_component* components,
                                       uint32_t n_components, const double* sigma_m2, uint32_t n_species) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        if (h->n_dom < 2) fail(ICARION_ERR_STATE, "icarion_gpu_set_domains with two or more domains not called");
        if (domain >= h->n_dom) fail(ICARION_ERR_INVALID, "domain index out of range");
        if (n_components == 0) {
            h->dom_mix[domain].clear();
            h->dom_mix_sig[domain].clear();
            h->dom_mix_ns[domain] = 0;
            h->dom_mix_all[domain].clear();
            h->dom_emix_sig[domain].clear();
            return;
        }
        if (!components || !sigma_m2) fail(ICARION_ERR_INVALID, "NULL argument");
        if (!h->have_species || n_species != h->n_species)
            fail(ICARION_ERR_INVALID, "n_species must match the species table (set_species first)");
        build_hss_mixture(components, n_components, sigma_m2, n_species, h->dom_mix[domain], h->dom_mix_sig[domain]);
        h->dom_mix_ns[domain] = n_species;
        h->dom_mix_all[domain].assign(components, components + n_components);
        h->dom_emix_sig[domain].assign(sigma_m2, sigma_m2 + static_cast<size_t>(n_species) * n_components);
    });
}

int icarion_gpu_set_domain_species(icarion_gpu* h, uint32_t domain, const icarion_species_desc* species,
                                   uint32_t n_species) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        if (h->n_dom < 2) fail(ICARION_ERR_STATE, "icarion_gpu_set_domains with two or more domains not called");
        if (domain >= h->n_dom) fail(ICARION_ERR_INVALID, "domain index out of range");
        if (!species || n_species == 0) {
            h->dom_species[domain].clear();
            return;
        }
        if (!h->have_species || n_species != h->n_species)
            fail(ICARION_ERR_INVALID, "n_species must match the species table (set_species first)");
        for (uint32_t s = 0; s < n_species; ++s)
            if (species[s].atom_offset != h->species_host[s].atom_offset ||
                species[s].atom_count != h->species_host[s].atom_count)
                fail(ICARION_ERR_INVALID, "species " + std::to_string(s) +
                                              ": a domain's row keeps the handle's atom range (one geometry per species)");
        h->dom_species[domain] = species_rows(species, n_species, h->n_atoms);
    });
}

int icarion_gpu_set_domain_waveforms(icarion_gpu* h, uint32_t domain, const icarion_waveform_desc* const* slots) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        if (h->n_dom < 2) fail(ICARION_ERR_STATE, "icarion_gpu_set_domains with two or more domains not called");
        if (domain >= h->n_dom) fail(ICARION_ERR_INVALID, "domain index out of range");
        std::unique_ptr<HostWaveform> w[ICARION_N_VOLTAGE_SLOTS];
        parse_waveform_slots(slots, w);
        for (int k = 0; k < ICARION_N_VOLTAGE_SLOTS; ++k) h->dom_wave[domain][k] = std::move(w[k]);
    });
}

int icarion_gpu_set_pressure_waveform(icarion_gpu* h, const icarion_waveform_desc* pressure) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        h->pressure = pressure ? parse_waveform(pressure) : nullptr;
    });
}

int icarion_gpu_set_gas_flow(icarion_gpu* h, int32_t flow_model, double axial_max_velocity_m_s) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        if (flow_model != ICARION_FLOW_CONSTANT && flow_model != ICARION_FLOW_AXIAL_PARABOLIC)
            fail(ICARION_ERR_INVALID, "unknown flow model");
        if (!std::isfinite(axial_max_velocity_m_s)) fail(ICARION_ERR_INVALID, "the flow velocity must be finite");
        h->flow_model = flow_model;
        h->flow_umax = flow_model == ICARION_FLOW_CONSTANT ? 0.0 : axial_max_velocity_m_s;
    });
}

int icarion_gpu_set_domain_gas_flow(icarion_gpu* h, uint32_t domain, int32_t flow_model,
                                    double axial_max_velocity_m_s) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        if (h->n_dom < 2) fail(ICARION_ERR_STATE, "icarion_gpu_set_domains with two or more domains not called");
        if (domain >= h->n_dom) fail(ICARION_ERR_INVALID, "domain index out of range");
        if (flow_model != ICARION_FLOW_CONSTANT && flow_model != ICARION_FLOW_AXIAL_PARABOLIC)
            fail(ICARION_ERR_INVALID, "unknown flow model");
        if (!std::isfinite(axial_max_velocity_m_s)) fail(ICARION_ERR_INVALID, "the flow velocity must be finite");
        h->dom_flow[domain] = {flow_model, flow_model == ICARION_FLOW_CONSTANT ? 0.0 : axial_max_velocity_m_s};
    });
}

int icarion_gpu_set_domain_pressure_waveform(icarion_gpu* h, uint32_t domain, const icarion_waveform_desc* pressure) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        if (h->n_dom < 2) fail(ICARION_ERR_STATE, "icarion_gpu_set_domains with two or more domains not called");
        if (domain >= h->n_dom) fail(ICARION_ERR_INVALID, "domain index out of range");
        h->dom_pressure[domain] = pressure ? parse_waveform(pressure) : nullptr;
    });
}

int icarion_gpu_set_ehss_offline(icarion_gpu* h, const uint32_t* n_orient, const uint32_t* mu_stride,
                                 const double* sigma_eff, const double* mu, uint32_t n_species) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        if (n_species == 0) {
            h->off_species = 0;
            h->off_tab.release();
            h->off_sigma.release();
            h->off_mu.release();
            return;
        }
        if (!n_orient || !mu_stride) fail(ICARION_ERR_INVALID, "NULL argument");
        if (!h->have_species || n_species != h->n_species)
            fail(ICARION_ERR_INVALID, "n_species must match the species table (set_species first)");
        std::vector<uint4> tab(n_species);
        uint64_t n_sig = 0, n_mu = 0;
        for (uint32_t sp = 0; sp < n_species; ++sp) {
            if (n_orient[sp] > 0 && mu_stride[sp] == 0)
                fail(ICARION_ERR_INVALID, "species " + std::to_string(sp) + ": mu_stride must be > 0");
            tab[sp] = make_uint4(static_cast<uint32_t>(n_sig), n_orient[sp], n_orient[sp] ? mu_stride[sp] : 0u,
                                 static_cast<uint32_t>(n_mu));
            n_sig += n_orient[sp];
            n_mu += static_cast<uint64_t>(n_orient[sp]) * (n_orient[sp] ? mu_stride[sp] : 0u);
            if (n_sig > 0xFFFFFFFFull || n_mu > 0xFFFFFFFFull)
                fail(ICARION_ERR_UNSUPPORTED, "offline EHSS sets above 2^32 entries");
        }
        if (n_sig == 0) fail(ICARION_ERR_INVALID, "no species has an offline set");
        if (!sigma_eff || !mu) fail(ICARION_ERR_INVALID, "NULL argument");
        for (uint64_t i = 0; i < n_sig; ++i)
            if (!std::isfinite(sigma_eff[i])) fail(ICARION_ERR_INVALID, "sigma_eff values must be finite");
        for (uint64_t i = 0; i < n_mu; ++i)
            if (!std::isfinite(mu[i])) fail(ICARION_ERR_INVALID, "mu values must be finite");
        set_device(h);
        copy_in(h->off_tab, tab.data(), tab.size(), h->stream, "upload offline EHSS index");
        copy_in(h->off_sigma, sigma_eff, static_cast<size_t>(n_sig), h->stream, "upload offline sigma_eff");
        copy_in(h->off_mu, mu, static_cast<size_t>(n_mu), h->stream, "upload offline mu");
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        h->off_species = n_species;
    });
}

// CollisionGeometry::quaternion_to_rotation (CollisionGeometry.cpp:73-104):
// normalise, then the rotation matrix, row-major.
void quaternion_rotation(const double q[4], double R[9]) {
    double w = q[0], x = q[1], y = q[2], z = q[3];
    const double norm = std::sqrt(w * w + x * x + y * y + z * z);
    if (norm <= 0.0) {
        const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        std::memcpy(R, I, sizeof(I));
        return;
    }
    w /= norm;
    x /= norm;
    y /= norm;
    z /= norm;
    R[0] = 1.0 - 2.0 * (y * y + z * z);
    R[1] = 2.0 * (x * y - z * w);
    R[2] = 2.0 * (x * z + y * w);
    R[3] = 2.0 * (x * y + z * w);
    R[4] = 1.0 - 2.0 * (x * x + z * z);
    R[5] = 2.0 * (y * z - x * w);
    R[6] = 2.0 * (x * z - y * w);
    R[7] = 2.0 * (y * z + x * w);
    R[8] = 1.0 - 2.0 * (x * x + y * y);
}

int icarion_gpu_set_ehss_orientations(icarion_gpu* h, const uint32_t* n_orient, const double* quat_wxyz,
                                      const double* area_m2, uint32_t n_species) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        if (n_species == 0) {
            h->ori_species = 0;
            h->ori_tab.release();
            h->ori_R.release();
            h->ori_area.release();
            return;
        }
        if (!n_orient) fail(ICARION_ERR_INVALID, "NULL argument");
        if (!h->have_species || n_species != h->n_species)
            fail(ICARION_ERR_INVALID, "n_species must match the species table (set_species first)");
        std::vector<uint2> tab(n_species);
        uint64_t total = 0;
        for (uint32_t sp = 0; sp < n_species; ++sp) {
            tab[sp] = make_uint2(static_cast<uint32_t>(total), n_orient[sp]);
            total += n_orient[sp];
            if (total > 0xFFFFFFFFull / 9) fail(ICARION_ERR_UNSUPPORTED, "too many EHSS orientation samples");
        }
        if (total == 0) fail(ICARION_ERR_INVALID, "no species has orientation samples");
        if (!quat_wxyz || !area_m2) fail(ICARION_ERR_INVALID, "NULL argument");
        std::vector<double> R(9 * total);
        for (uint64_t i = 0; i < total; ++i) {
            for (int c = 0; c < 4; ++c)
                if (!std::isfinite(quat_wxyz[4 * i + c])) fail(ICARION_ERR_INVALID, "quaternions must be finite");
            if (!std::isfinite(area_m2[i])) fail(ICARION_ERR_INVALID, "areas must be finite");
            quaternion_rotation(quat_wxyz + 4 * i, R.data() + 9 * i);
        }
        set_device(h);
        copy_in(h->ori_tab, tab.data(), tab.size(), h->stream, "upload EHSS orientation index");
        copy_in(h->ori_R, R.data(), R.size(), h->stream, "upload EHSS orientations");
        copy_in(h->ori_area, area_m2, static_cast<size_t>(total), h->stream, "upload EHSS orientation areas");
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        h->ori_species = n_species;
    });
}

int icarion_gpu_upload_field_terms(icarion_gpu* h, uint32_t nx, uint32_t ny, uint32_t nz,
                                   const double* xs, const double* ys, const double* zs,
                                   const icarion_field_term_desc* terms, uint32_t n_terms) {
    return guarded([&] {
        if (!h || !xs || !ys || !zs || !terms) fail(ICARION_ERR_INVALID, "NULL argument");
        if (nx < 2 || ny < 2 || nz < 2) fail(ICARION_ERR_INVALID, "field array needs >= 2 nodes per axis");
        if (n_terms == 0) fail(ICARION_ERR_INVALID, "no field terms");
        if (n_terms > static_cast<uint32_t>(kMaxFieldTerms))
            fail(ICARION_ERR_UNSUPPORTED, "at most " + std::to_string(kMaxFieldTerms) + " field terms");
        const double* axes[3] = {xs, ys, zs};
        const uint32_t dims[3] = {nx, ny, nz};
        for (int k = 0; k < 3; ++k) {
            for (uint32_t i = 0; i < dims[k]; ++i)
                if (!std::isfinite(axes[k][i])) fail(ICARION_ERR_INVALID, "axis values must be finite");
            for (uint32_t i = 0; i + 1 < dims[k]; ++i)
                if (axes[k][i + 1] < axes[k][i]) fail(ICARION_ERR_INVALID, "axes must be non-decreasing");
            if (!(axes[k][dims[k] - 1] > axes[k][0])) fail(ICARION_ERR_INVALID, "an axis has zero extent");
        }
        for (uint32_t j = 0; j < n_terms; ++j) {
            if (!terms[j].E) fail(ICARION_ERR_INVALID, "field term without data");
            if (terms[j].dtype != 0 && terms[j].dtype != 1)
                fail(ICARION_ERR_INVALID, "dtype must be 0 (float) or 1 (double)");
            if (terms[j].scale_kind < ICARION_FIELD_SCALE_CONSTANT || terms[j].scale_kind > ICARION_FIELD_SCALE_RF)
                fail(ICARION_ERR_INVALID, "unknown field scale kind");
        }
        set_device(h);
        upload_field_data(h, static_cast<size_t>(nx) * ny * nz, terms, n_terms);
        icarion_gpu::FieldSrc f;
        f.nx = static_cast<int>(nx); f.ny = static_cast<int>(ny); f.nz = static_cast<int>(nz);
        f.n_terms = static_cast<int>(n_terms);
        double d[3];
        bool uniform = true;
        for (int k = 0; k < 3; ++k) uniform &= uniform_axis(axes[k], dims[k], d[k]);
        f.nonuniform = !uniform;
        f.x0 = xs[0]; f.y0 = ys[0]; f.z0 = zs[0];
        f.dx = uniform ? d[0] : 1.0; f.dy = uniform ? d[1] : 1.0; f.dz = uniform ? d[2] : 1.0;
        f.x_last = xs[nx - 1]; f.y_last = ys[ny - 1]; f.z_last = zs[nz - 1];
        for (auto& a : h->ax64) a.release();
        for (auto& a : h->ax32) a.release();
        if (!uniform) {
            for (int k = 0; k < 3; ++k) {
                std::vector<float> af(axes[k], axes[k] + dims[k]);
                copy_in(h->ax64[k], axes[k], dims[k], h->stream, "upload axis");
                copy_in(h->ax32[k], af.data(), dims[k], h->stream, "upload axis (float)");
                hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
            }
        }
        for (uint32_t j = 0; j < n_terms; ++j) {
            f.kind[j] = terms[j].scale_kind;
            f.cst[j] = terms[j].constant_scale;
            f.freq[j] = terms[j].frequency_Hz;
            f.phase[j] = terms[j].phase_rad;
        }
        h->fsrc = f;
        h->have_field = true;
    });
}

int icarion_gpu_upload_state(icarion_gpu* h, const icarion_ion_soa* ions) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        do_upload(h, ions);
    });
}

int icarion_gpu_step(icarion_gpu* h, double t0, double dt, uint64_t n_steps, uint64_t step_offset,
                     icarion_step_stats* stats) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        do_step(h, t0, dt, n_steps, step_offset, stats);
    });
}

int icarion_gpu_step_noalloc(icarion_gpu* h, double t0, double dt, uint64_t n_steps, uint64_t step_offset,
                             icarion_step_stats* stats) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        if (!h->scratch_dev) {
            std::lock_guard<std::mutex> lock(g_scratch_mu);
            fail(ICARION_ERR_HIP, h->device < kMaxScratchDevices && !g_scratch_error[h->device].empty()
                                      ? g_scratch_error[h->device]
                                      : std::string("no scratch array on devices above 63"));
        }
        Workspace ws;
        ws.counters = reinterpret_cast<unsigned long long*>(h->scratch_dev);
        ws.table = reinterpret_cast<StepEntry*>(h->scratch_dev + kNoAllocCounterBytes);
        ws.max_steps = ICARION_NOALLOC_MAX_STEPS;
        do_step(h, t0, dt, n_steps, step_offset, stats, nullptr, &ws);
    });
}

int icarion_gpu_step_snapshots(icarion_gpu* h, double t0, double dt, uint64_t n_steps, uint64_t step_offset,
                               double write_interval_s, int write_first, double* next_write_s,
                               icarion_snapshot_fn fn, void* user, icarion_step_stats* stats) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "handle is NULL");
        if (!fn || !next_write_s) fail(ICARION_ERR_INVALID, "fn and next_write_s are required");
        if (!(write_interval_s > 0.0) || !std::isfinite(write_interval_s))
            fail(ICARION_ERR_INVALID, "write_interval_s must be finite and > 0");
        if (std::isnan(*next_write_s)) fail(ICARION_ERR_INVALID, "*next_write_s is NaN");
        const SnapshotReq r{write_interval_s, write_first != 0, next_write_s, fn, user};
        do_step(h, t0, dt, n_steps, step_offset, stats, &r);
    });
}

int icarion_gpu_download_state(icarion_gpu* h, icarion_ion_soa* ions) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        do_download(h, ions);
    });
}


int icarion_gpu_synth_clusters(icarion_gpu* h, uint32_t n_clusters, const uint32_t* counts, double r_min_m,
                               double r_max_m, uint64_t seed, float* atoms_out, double* sigma_out,
                               float* features_out) {
    return guarded([&] {
        if (!h || !counts || n_clusters == 0) fail(ICARION_ERR_INVALID, "no clusters requested");
        if (!h->have_domain) fail(ICARION_ERR_STATE, "icarion_gpu_set_domain must precede synth_clusters");
        if (!(r_min_m > 0.0 && r_max_m >= r_min_m)) fail(ICARION_ERR_INVALID, "atom radii must satisfy 0 < r_min <= r_max");
        std::vector<uint32_t> offs(n_clusters);
        uint64_t total = 0;
        for (uint32_t c = 0; c < n_clusters; ++c) {
            if (counts[c] == 0) fail(ICARION_ERR_INVALID, "a cluster needs at least one atom");
            offs[c] = static_cast<uint32_t>(total);
            total += counts[c];
            if (total > 0xFFFFFFFFull) fail(ICARION_ERR_INVALID, "at most 2^32-1 atoms in total");
        }
        set_device(h);
        hipStream_t s = h->stream;
        h->atoms_f32.resize(total);
        copy_in(h->atom_offsets, offs.data(), n_clusters, s, "upload offsets");
        copy_in(h->atom_counts, counts, n_clusters, s, "upload counts");
        hip_check(launch_synth_clusters(h->atoms_f32.p, h->atom_offsets.p, h->atom_counts.p, n_clusters,
                                        static_cast<float>(r_min_m), static_cast<float>(r_max_m), seed, s),
                  "synth_clusters");
        const std::vector<double> rot = oapa_rotations();
        DevBuf<double> rot_d, sig_d;
        DevBuf<float> feat_d;
        copy_in(rot_d, rot.data(), rot.size(), s, "upload rotations");
        sig_d.resize(n_clusters);
        feat_d.resize(3ull * n_clusters);
        const double r_gas = h->dom.gas_radius_m;
        hip_check(launch_oapa(h->atoms_f32.p, h->atom_offsets.p, h->atom_counts.p, n_clusters, rot_d.p, r_gas,
                              sig_d.p, s), "oapa");
        hip_check(launch_features(h->atoms_f32.p, h->atom_offsets.p, h->atom_counts.p, n_clusters, r_gas,
                                  sig_d.p, feat_d.p, s), "features");
        std::vector<double> sig(n_clusters);
        std::vector<float> feat(3ull * n_clusters);
        copy_out(sig.data(), sig_d, n_clusters, s, "download sigma");
        copy_out(feat.data(), feat_d, 3ull * n_clusters, s, "download features");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        std::vector<SpeciesDev> sd(n_clusters);
        std::vector<icarion_species_desc> sh(n_clusters);
        for (uint32_t c = 0; c < n_clusters; ++c) {
            sd[c] = SpeciesDev{};
            sd[c].sigma_ehss = sig[c];
            sd[c].atom_offset = offs[c];
            sd[c].atom_count = counts[c];
            for (int f = 0; f < 3; ++f) sd[c].feat[f] = feat[3ull * c + f];
            sh[c] = icarion_species_desc{};
            sh[c].sigma_ehss_m2 = sig[c];
            sh[c].atom_offset = offs[c];
            sh[c].atom_count = counts[c];
            for (int f = 0; f < 3; ++f) sh[c].surrogate_features[f] = feat[3ull * c + f];
        }
        copy_in(h->species, sd.data(), n_clusters, s, "upload species");
        if (atoms_out) copy_out(reinterpret_cast<float4*>(atoms_out), h->atoms_f32, total, s, "download atoms");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (sigma_out) std::memcpy(sigma_out, sig.data(), sizeof(double) * n_clusters);
        if (features_out) std::memcpy(features_out, feat.data(), sizeof(float) * 3 * n_clusters);
        h->species_host = std::move(sh);
        h->n_species = n_clusters;
        h->n_atoms = static_cast<uint32_t>(total);
        h->have_species = true;
    });
}

int icarion_gpu_ehss_precompute(icarion_gpu* h, uint32_t cluster_begin, uint32_t n_clusters, uint32_t n_orient,
                                uint32_t n_sigma, uint32_t n_mu, uint32_t max_tries, uint64_t seed,
                                double* sigma_out, double* mu_out) {
    return guarded([&] {
        if (!h || !sigma_out || (n_mu && !mu_out)) fail(ICARION_ERR_INVALID, "NULL argument");
        if (!h->have_domain || !h->have_species || !h->atoms_f32.p)
            fail(ICARION_ERR_STATE, "needs a domain and a float atom table");
        if (static_cast<uint64_t>(cluster_begin) + n_clusters > h->n_species)
            fail(ICARION_ERR_INVALID, "cluster range exceeds the species table");
        if (n_orient == 0) fail(ICARION_ERR_INVALID, "n_orient must be > 0");
        set_device(h);
        hipStream_t s = h->stream;
        const uint64_t nt = static_cast<uint64_t>(n_clusters) * n_orient;
        DevBuf<double> sig, mu;
        sig.resize(nt);
        mu.resize(std::max<uint64_t>(nt * n_mu, 1));
        hip_check(launch_precompute(h->atoms_f32.p, h->atom_offsets.p, h->atom_counts.p, cluster_begin, n_clusters,
                                    n_orient, n_sigma, n_mu, max_tries, seed, h->dom.gas_radius_m, sig.p, mu.p, s),
                  "precompute");
        copy_out(sigma_out, sig, nt, s, "download sigma");
        if (n_mu) copy_out(mu_out, mu, nt * n_mu, s, "download mu");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

int icarion_gpu_surrogate_set_params(icarion_gpu* h, const float* params, uint32_t n_params) {
    return guarded([&] {
        if (!h || !params) fail(ICARION_ERR_INVALID, "NULL argument");
        if (n_params != static_cast<uint32_t>(kMlpParams)) fail(ICARION_ERR_INVALID, "wrong parameter count");
        install_params(h, params);
    });
}

int icarion_gpu_surrogate_get_params(icarion_gpu* h, float* params, uint32_t n_params) {
    return guarded([&] {
        if (!h || !params) fail(ICARION_ERR_INVALID, "NULL argument");
        if (n_params != static_cast<uint32_t>(kMlpParams)) fail(ICARION_ERR_INVALID, "wrong parameter count");
        if (!h->have_surrogate) fail(ICARION_ERR_STATE, "no surrogate parameters");
        std::memcpy(params, h->params.data(), sizeof(float) * kMlpParams);
    });
}

int icarion_gpu_surrogate_train(icarion_gpu* h, const float* X, const float* T, size_t n,
                                const icarion_train_desc* desc, float* loss_per_epoch) {
    return guarded([&] {
        if (!h || !X || !T || !desc) fail(ICARION_ERR_INVALID, "NULL argument");
        const int B = static_cast<int>(desc->batch);
        if (B <= 0 || static_cast<size_t>(B) > n) fail(ICARION_ERR_INVALID, "batch must be in [1, n]");
        if (n > 0xFFFFFFFFull) fail(ICARION_ERR_INVALID, "at most 2^32-1 samples");
        set_device(h);
        hipStream_t s = h->stream;
        if (!h->have_surrogate) {
            const std::vector<float> p0 = init_params(desc->seed);
            h->params = p0;
        }
        DevBuf<float> Xd, Td;
        copy_in(Xd, X, n * kMlpIn, s, "upload X");
        copy_in(Td, T, n * kMlpOut, s, "upload T");
        TrainDev td;
        td.alloc(B);
        copy_in(td.params, h->params.data(), kMlpParams, s, "upload params");
        hipLaunchKernelGGL(zero_f32_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, td.m.p, kMlpParams);
        hipLaunchKernelGGL(zero_f32_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, td.v.p, kMlpParams);
        DevBuf<uint32_t> ctr, perm_d, keys_a, keys_b, idx;
        DevBuf<unsigned char> sort_temp;
        ctr.resize(2);
        hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, ctr.p, 2);
        const size_t n_batches = n / static_cast<size_t>(B);
        // the epoch's sample order: a device permutation of all n samples, of
        // which the first n_batches * B are used
        perm_d.resize(n);
        keys_a.resize(n);
        keys_b.resize(n);
        idx.resize(n);
        const size_t temp_bytes = brick_sort_temp_bytes(static_cast<uint32_t>(n));
        sort_temp.resize(temp_bytes);
        const TrainBuffers tb = td.view(td.Xb.p, td.Tb.p, B);
        // one training step, captured once as a HIP graph and replayed: the batch
        // offset and the Adam step count live in device counters
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipLaunchKernelGGL(gather_batch_kernel, dim3((B + 255) / 256), dim3(256), 0, s, Xd.p, Td.p, perm_d.p, ctr.p,
                           td.Xb.p, td.Tb.p, B);
        hipError_t ce = mlp_fwd_bwd(tb, s);
        hipLaunchKernelGGL(adam_graph_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, td.params.p, td.grad.p,
                           td.m.p, td.v.p, kMlpParams, desc->lr, desc->beta1, desc->beta2, desc->eps, ctr.p);
        hipLaunchKernelGGL(advance_counters_kernel, dim3(1), dim3(64), 0, s, ctr.p);
        const hipError_t ee = hipStreamEndCapture(s, &graph);
        hip_check(ce, "capture training step");
        hip_check(ee, "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        try {
            for (uint32_t ep = 0; ep < desc->epochs; ++ep) {
                hip_check(random_permutation(static_cast<uint32_t>(n), desc->seed, ep, keys_a.p, keys_b.p, idx.p,
                                             perm_d.p, sort_temp.p, temp_bytes, s),
                          "epoch permutation");
                hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, s, td.loss.p, 1);
                hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, ctr.p, 1);  // batch index only
                for (size_t b = 0; b < n_batches; ++b) hip_check(hipGraphLaunch(exec, s), "hipGraphLaunch");
                float l = 0.f;
                hip_check(hipMemcpyAsync(&l, td.loss.p, sizeof(float), hipMemcpyDeviceToHost, s), "read loss");
                hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
                if (loss_per_epoch) loss_per_epoch[ep] = l / static_cast<float>(n_batches);
            }
        } catch (...) {
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            throw;
        }
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        std::vector<float> p(kMlpParams);
        copy_out(p.data(), td.params, kMlpParams, s, "download params");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        td.release();
        install_params(h, p.data());
    });
}

int icarion_gpu_surrogate_grad(icarion_gpu* h, const float* X, const float* T, size_t n, float* grad_out,
                               float* loss_out) {
    return guarded([&] {
        if (!h || !X || !T || !grad_out || n == 0) fail(ICARION_ERR_INVALID, "NULL argument");
        if (!h->have_surrogate) fail(ICARION_ERR_STATE, "no surrogate parameters");
        if (n > (1u << 24)) fail(ICARION_ERR_INVALID, "at most 2^24 samples per gradient call");
        set_device(h);
        hipStream_t s = h->stream;
        const int B = static_cast<int>(n);
        TrainDev td;
        td.alloc(B);
        copy_in(td.params, h->params.data(), kMlpParams, s, "upload params");
        copy_in(td.Xb, X, n * kMlpIn, s, "upload X");
        copy_in(td.Tb, T, n * kMlpOut, s, "upload T");
        hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, s, td.loss.p, 1);
        hip_check(mlp_fwd_bwd(td.view(td.Xb.p, td.Tb.p, B), s), "fwd/bwd");
        copy_out(grad_out, td.grad, kMlpParams, s, "download grad");
        float l = 0.f;
        hip_check(hipMemcpyAsync(&l, td.loss.p, sizeof(float), hipMemcpyDeviceToHost, s), "read loss");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (loss_out) *loss_out = l;
        td.release();
    });
}

int icarion_gpu_surrogate_infer(icarion_gpu* h, const float* X, size_t n, float* Y) {
    return guarded([&] {
        if (!h || !X || !Y) fail(ICARION_ERR_INVALID, "NULL argument");
        if (!h->have_surrogate) fail(ICARION_ERR_STATE, "no surrogate parameters");
        if (n == 0) return;
        if (n > 0xFFFFFFFFull) fail(ICARION_ERR_INVALID, "too many samples");
        set_device(h);
        hipStream_t s = h->stream;
        DevBuf<float> Xd, Yd;
        copy_in(Xd, X, n * kMlpIn, s, "upload X");
        Yd.resize(n * kMlpOut);
        hip_check(mlp_infer(h->mlp_blob.p, Xd.p, Yd.p, static_cast<uint32_t>(n), s), "mlp_infer");
        copy_out(Y, Yd, n * kMlpOut, s, "download Y");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}


int icarion_gpu_synchronize(icarion_gpu* h) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        set_device(h);
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

int icarion_gpu_batch_upload(icarion_gpu* h, const icarion_cloud_desc* clouds, uint32_t n_clouds) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        do_batch_upload(h, clouds, n_clouds);
    });
}

int icarion_gpu_batch_run(icarion_gpu* h, const icarion_batch_tuning* fp64, const icarion_batch_tuning* fp32,
                          icarion_cloud_stats* per_cloud, double* device_ms) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        do_batch_run(h, fp64, fp32, per_cloud, device_ms);
    });
}

int icarion_gpu_batch_autotune(icarion_gpu* h, uint32_t trial_steps, icarion_batch_tuning* best64,
                               icarion_batch_tuning* best32, double* rate64, double* rate32) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        do_batch_autotune(h, trial_steps, best64, best32, rate64, rate32);
    });
}

int icarion_gpu_batch_download(icarion_gpu* h, const icarion_cloud_desc* clouds, uint32_t n_clouds) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        do_batch_download(h, clouds, n_clouds);
    });
}

int icarion_gpu_step_batch(icarion_gpu* h, icarion_ion_soa* ions, double t, double dt,
                           uint64_t step_index, icarion_step_stats* stats) {
    return guarded([&] {
        if (!h) fail(ICARION_ERR_INVALID, "NULL handle");
        do_upload(h, ions);
        do_step(h, t, dt, 1, step_index, stats);
        do_download(h, ions);
    });
}

}  // extern "C"
