// icarion_config.hpp - C++17 helpers that turn the reference's evaluated
// configuration into this library's descriptors (include/icarion_gpu.h), so that
// `simulation.enable_gpu` runs the same JSON config schema on the device.
//
// The reference parses its JSON into DomainConfig / EnvironmentConfig /
// SimulationConfig / PhysicsConfig / SpeciesDatabase / Molecule (jsoncpp).  The
// structs below mirror the fields of those types that the batch ion-step path
// reads, with the reference's names and units, so a reference-side adaptor fills
// them by copying fields; nothing here parses JSON or includes a reference
// header.  The helpers restate the reference's own derivations:
//
//   Environment::derive          EnvironmentConfig::compute_derived_properties
//                                (src/core/config/types/EnvironmentConfig.h:150-282)
//   axial_V_from_EN              DomainConfigLoader::load_dc_fields
//                                (src/core/config/loader/DomainConfigLoader.cpp:413-440)
//   to_desc / to_physics         DomainConfig + SimulationConfig + PhysicsConfig
//                                -> icarion_domain_desc / icarion_physics_desc
//   hss_sigma                    HSSCollisionHandler single-gas sigma priority
//                                (src/core/physics/collisions/HSSCollisionHandler.cpp:263-310)
//   hss_mixture_sigma            the same per mixture component (HSSCollisionHandler.cpp:136-172)
//   ehss_sigma                   EHSSCollisionHandler::compute_effective_ccs
//                                (src/core/physics/collisions/EHSSCollisionHandler.cpp:458-547)
//   derive_ccs_for_target_gas    HSSCollisionHandler.cpp:419-447 / EHSSCollisionHandler.cpp:571-600
//   atom_table                   convert_molecule_to_geometry
//                                (src/core/physics/collisions/geometryUtils.h:35-73)
//   fused_step_eligible          the engine's batch-collision preconditions
//                                (src/core/integrator/SimulationEngineCollisionRuntime.cpp:26-63)
//                                and batch-integration precondition
//                                (src/core/integrator/SimulationEngineIntegration.cpp:23-45)
//
// Errors: a configuration the device path cannot take throws
// icarion::config::Unsupported (the adaptor then keeps the CPU path); malformed
// values throw std::invalid_argument.
#ifndef ICARION_CONFIG_HPP
#define ICARION_CONFIG_HPP

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <map>
#include <optional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "icarion_gpu.h"

namespace icarion {
namespace config {

// ---- constants (src/utils/constants.h:13-65) --------------------------------
constexpr double kAmuToKg = 1.66053906660e-27;
constexpr double kElemChargeC = 1.602176634e-19;
constexpr double kBoltzmann = 1.380649e-23;
constexpr double kAngstrom2ToM2 = 1e-20;
constexpr double kAngstromToM = 1e-10;
constexpr double kPi = 3.14159265358979323846;

class Unsupported : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

// Neutral gas properties: molar mass, hard-sphere radius, polarizability.
struct GasProps {
    double mass_kg = 0.0;
    double radius_m = 0.0;
    double polarizability_m3 = 0.0;
};

// The seven tabulated gases (constants.h:41-65).  false for any other name.
inline bool tabulated_gas(const std::string& name, GasProps& g) {
    struct Row {
        const char* name;
        double amu, radius, alpha;
    };
    static const Row rows[] = {
        {"N2", 28.0134, 1.82e-10, 1.74e-30},  {"He", 4.002602, 1.3e-10, 0.195e-30},
        {"Ar", 39.948, 1.70e-10, 1.596e-30},  {"O2", 31.9988, 1.73e-10, 1.567e-30},
        {"CO2", 44.0095, 1.65e-10, 2.612e-30}, {"Ne", 20.1797, 1.38e-10, 0.352e-30},
        {"H2O", 18.01528, 1.58e-10, 1.45e-30},
    };
    for (const Row& r : rows) {
        if (name == r.name) {
            g = GasProps{r.amu * kAmuToKg, r.radius, r.alpha};
            return true;
        }
    }
    return false;
}

// A single buffer gas by name, as compute_derived_properties' single-gas branch
// (EnvironmentConfig.h:229-270): the seven gases, "Air" as the N2/O2/Ar
// mole-fraction mean, "Default" or "" as He.
inline GasProps single_gas(const std::string& name) {
    GasProps g;
    if (tabulated_gas(name, g)) return g;
    if (name == "Air") {
        GasProps n2, o2, ar;
        tabulated_gas("N2", n2);
        tabulated_gas("O2", o2);
        tabulated_gas("Ar", ar);
        g.mass_kg = 0.78084 * n2.mass_kg + 0.20946 * o2.mass_kg + 0.00934 * ar.mass_kg;
        g.radius_m = 0.78084 * n2.radius_m + 0.20946 * o2.radius_m + 0.00934 * ar.radius_m;
        g.polarizability_m3 =
            0.78084 * n2.polarizability_m3 + 0.20946 * o2.polarizability_m3 + 0.00934 * ar.polarizability_m3;
        return g;
    }
    if (name == "Default" || name.empty()) {
        tabulated_gas("He", g);
        return g;
    }
    throw Unsupported("unknown gas_species '" + name + "'");
}

// ---- environment (EnvironmentConfig) ----------------------------------------
struct GasComponent {                 // GasMixtureComponent
    std::string species;
    double mole_fraction = 0.0;
    double cross_section_m2 = -1.0;   // < 0: not set
    double polarizability_m3 = -1.0;  // < 0: the gas table's
    bool participates_in_collisions = true;
    // derived
    double mass_kg = 0.0, radius_m = 0.0, density_m3 = 0.0, polarizability_m3_derived = 0.0;
};

struct Environment {
    double temperature_K = 300.0;
    double pressure_Pa = 101325.0;
    std::string gas_species = "He";
    std::array<double, 3> gas_velocity_m_s{{0.0, 0.0, 0.0}};
    // flow_model and its parameters (EnvironmentConfig.h:22-26, 60-62; names as
    // parse_flow_model accepts them, EnvironmentConfig.h:81-95)
    std::string flow_model = "constant";
    double axial_flow_velocity_m_s = 0.0;
    double axial_flow_max_velocity_m_s = 0.0;
    std::vector<GasComponent> gas_mixture;
    // derived
    double particle_density_m_3 = 0.0;
    double gas_mass_kg = 0.0;
    double gas_radius_m = 0.0;
    double gas_polarizability_m3 = 0.0;

    // compute_derived_properties: n = P/(kB T); a mixture's components get
    // n_i = (x_i / sum x) n and their gas-table values (an unknown species
    // takes He's), the mixture means weight them by fraction; a single gas takes
    // its table values.
    void derive() {
        if (!(temperature_K > 0.0) || !(pressure_Pa >= 0.0))
            throw std::invalid_argument("temperature_K must be > 0 and pressure_Pa >= 0");
        particle_density_m_3 = pressure_Pa / (kBoltzmann * temperature_K);
        if (!gas_mixture.empty()) {
            double frac_sum = 0.0;
            for (const GasComponent& c : gas_mixture) frac_sum += c.mole_fraction;
            const double norm = frac_sum > 0.0 ? frac_sum : 1.0;
            gas_mass_kg = gas_radius_m = gas_polarizability_m3 = 0.0;
            for (GasComponent& c : gas_mixture) {
                const double f = c.mole_fraction / norm;
                c.density_m3 = f * particle_density_m_3;
                GasProps g;
                if (!tabulated_gas(c.species, g)) tabulated_gas("He", g);
                c.mass_kg = g.mass_kg;
                c.radius_m = g.radius_m;
                c.polarizability_m3_derived = c.polarizability_m3 > 0.0 ? c.polarizability_m3 : g.polarizability_m3;
                gas_mass_kg += f * c.mass_kg;
                gas_radius_m += f * c.radius_m;
                gas_polarizability_m3 += f * c.polarizability_m3_derived;
            }
            return;
        }
        const GasProps g = single_gas(gas_species);
        gas_mass_kg = g.mass_kg;
        gas_radius_m = g.radius_m;
        gas_polarizability_m3 = g.polarizability_m3;
    }
};

// ---- domain (DomainConfig: geometry, fields, env) -----------------------------
struct Geometry {
    std::array<double, 3> origin_m{{0.0, 0.0, 0.0}};
    double length_m = 0.0;
    double radius_m = 0.0;
    double radius_in_m = 0.0;
    double radius_out_m = 0.0;
    double radius_char_m = 0.0;
    double end_aperture_m = 0.0;
    // DomainConfig::rotation_global_to_local / _local_to_global, row-major
    std::array<double, 9> rotation_global_to_local{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    std::array<double, 9> rotation_local_to_global{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
};

// The constant values of the domain's fields (a waveform-valued voltage goes to
// icarion_gpu_set_waveforms instead; its constant here is ignored).
struct Fields {
    std::optional<double> EN_Td;        // DC.EN_Td
    std::optional<double> axial_V;      // DC.axial_V (overrides EN_Td)
    std::optional<double> quad_V;       // DC.quad_V
    std::optional<double> radial_V;     // DC.radial_V
    std::optional<double> rf_voltage_V;     // RF.voltage_V
    std::optional<double> rf_frequency_Hz;  // RF.frequency_Hz
    double rf_phase_rad = 0.0;          // RF.phase_rad (not read by the analytic fields)
    bool magnetic_enabled = false;      // B / magnetic: enabled or field_strength_T given
    std::array<double, 3> field_strength_T{{0.0, 0.0, 0.0}};
    bool field_array = false;           // field_array_terms present (GridFieldProvider)
    double field_array_scale = 1.0;     // one term is used unscaled (PhysicsSetup.cpp:168-171)
};

struct Domain {
    std::string name;
    std::string instrument = "IMS";
    Geometry geometry;
    Environment env;
    Fields fields;
};

struct Simulation {                     // SimulationConfig
    std::string integrator = "RK4";
    double dt_s = 0.0;
    uint64_t rng_seed = 0;
    bool enable_gpu = false;
};

struct Physics {                        // PhysicsConfig
    std::string collision_model = "NoCollisions";
    bool collision_multi_event_mode = false;
    int collision_subcycles_per_step = 1;
    int collision_max_events_per_step = 1;
    bool ou_thermal_kicks = false;      // Langevin / Friction: OU kicks at the ion's gamma
};

inline std::string upper(std::string s) {
    for (char& c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
    return s;
}

inline std::string lower(std::string s) {
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

// InstrumentType names (src/core/config/types/InstrumentTypes.cpp:25-32) of the
// instruments the device fields cover.
inline int32_t instrument_code(const std::string& name) {
    const std::string k = upper(name);
    if (k == "IMS" || k == "SIFDT-MS" || k == "SIFDT_MS" || k == "SIFDT") return ICARION_INSTRUMENT_IMS;
    if (k == "TIMS" || k == "TIMS_TUNNEL") return ICARION_INSTRUMENT_TIMS;
    if (k == "QUADRUPOLE" || k == "QUADRUPOLE_RF" || k == "QUADRUPOLERF" || k == "SLIM")
        return ICARION_INSTRUMENT_QUADRUPOLE;
    if (k == "ORBITRAP") return ICARION_INSTRUMENT_ORBITRAP;
    throw Unsupported("instrument '" + name + "' has no device field");
}

// SolverType names (EnumMapper.h: "RK4", "RK45", "Boris").
inline int32_t integrator_code(const std::string& name) {
    const std::string k = upper(name);
    if (k == "RK4") return ICARION_INTEGRATOR_RK4;
    if (k == "BORIS") return ICARION_INTEGRATOR_BORIS;
    if (k == "RK45") return ICARION_INTEGRATOR_RK45;
    throw Unsupported("integrator '" + name + "' has no device path");
}

// CollisionModel names (EnumMapper::parse_collision_model, case-insensitive).
inline int32_t collision_code(const std::string& name) {
    const std::string k = lower(name);
    if (k == "none" || k == "nocollisions" || k == "vacuum") return ICARION_COLLISION_NONE;
    if (k == "hss" || k == "hsmc") return ICARION_COLLISION_HSS;
    if (k == "ehss") return ICARION_COLLISION_EHSS;
    if (k == "langevin") return ICARION_COLLISION_LANGEVIN;
    if (k == "friction") return ICARION_COLLISION_FRICTION;
    throw Unsupported("collision model '" + name + "' has no device path");
}

// E/N -> axial voltage: V = (E/N) 1e-21 N L with N = P/(kB T)
// (DomainConfigLoader.cpp:419-427).
inline double axial_V_from_EN(double EN_Td, double pressure_Pa, double temperature_K, double length_m) {
    const double k_B = 1.380649e-23;
    const double N = pressure_Pa / (k_B * temperature_K);
    const double E_Vm = EN_Td * 1e-21 * N;
    return E_Vm * length_m;
}

// The domain descriptor of a domain whose environment has been derived
// (Environment::derive).  detector_z_m is the arrival plane of the in-kernel
// first-crossing times (NaN: none; the reference has no such plane on the
// engine path, its analysis finds arrivals in the snapshots).
// FlowModelKind of a flow_model name: 0 Constant, 1 AxialUniform, 2 AxialParabolic
// (EnvironmentConfig::parse_flow_model, case-insensitive).
inline int flow_kind(const std::string& name) {
    std::string n = name;
    for (char& c : n) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    if (n == "constant" || n == "legacy") return 0;
    if (n == "axial_uniform" || n == "uniform" || n == "axial_const") return 1;
    if (n == "axial_parabolic" || n == "parabolic" || n == "poiseuille") return 2;
    throw std::invalid_argument("unknown flow_model '" + name + "'");
}

// Whether the device step covers the domain's flow model: Constant, and
// AxialUniform, whose gas velocity gas_velocity_at (EnvironmentConfig.h:97-129)
// gives as the constant R_l2g (0, 0, u) everywhere; AxialParabolic on a
// single-domain handle (icarion_gpu_set_gas_flow with gas_flow(e)), while
// several domains keep it on the CPU path.
inline bool flow_covered(const Environment& e, bool single_domain = true) {
    return flow_kind(e.flow_model) != 2 || single_domain;
}

// The flow profile icarion_gpu_set_gas_flow takes for the environment:
// AxialParabolic with axial_flow_max_velocity_m_s, else the constant model
// (AxialUniform goes into the descriptor's gas velocity, see to_desc).
struct GasFlow {
    int32_t model = ICARION_FLOW_CONSTANT;
    double axial_max_velocity_m_s = 0.0;
};

inline GasFlow gas_flow(const Environment& e) {
    GasFlow f;
    if (flow_kind(e.flow_model) == 2) {
        f.model = ICARION_FLOW_AXIAL_PARABOLIC;
        f.axial_max_velocity_m_s = e.axial_flow_max_velocity_m_s;
    }
    return f;
}

// gas_flow_applied: the caller passes gas_flow(d.env) to
// icarion_gpu_set_gas_flow, so an AxialParabolic environment is taken (the
// descriptor's gas velocity is then not read and holds 0); without it such an
// environment throws Unsupported.
inline icarion_domain_desc to_desc(const Domain& d, double detector_z_m = NAN, bool gas_flow_applied = false) {
    icarion_domain_desc o{};
    o.instrument = instrument_code(d.instrument);
    o.use_field_array = d.fields.field_array ? 1 : 0;
    o.magnetic_enabled = d.fields.magnetic_enabled ? 1 : 0;
    for (int k = 0; k < 3; ++k) o.origin_m[k] = d.geometry.origin_m[k];
    o.length_m = d.geometry.length_m;
    o.radius_m = d.geometry.radius_m;
    o.radius_in_m = d.geometry.radius_in_m;
    o.radius_out_m = d.geometry.radius_out_m;
    o.radius_char_m = d.geometry.radius_char_m;
    for (int k = 0; k < 9; ++k) {
        o.rot_g2l[k] = d.geometry.rotation_global_to_local[k];
        o.rot_l2g[k] = d.geometry.rotation_local_to_global[k];
    }
    const Environment& e = d.env;
    if (!(e.particle_density_m_3 > 0.0 || e.pressure_Pa == 0.0) || !(e.gas_mass_kg > 0.0))
        throw std::invalid_argument("domain '" + d.name + "': environment not derived (call Environment::derive)");
    // DC.axial_V overrides EN_Td; EN_Td converts with the domain's P, T and L
    if (d.fields.axial_V)
        o.dc_axial_V = *d.fields.axial_V;
    else if (d.fields.EN_Td)
        o.dc_axial_V = axial_V_from_EN(*d.fields.EN_Td, e.pressure_Pa, e.temperature_K, d.geometry.length_m);
    o.dc_quad_V = d.fields.quad_V.value_or(0.0);
    o.dc_radial_V = d.fields.radial_V.value_or(0.0);
    o.rf_voltage_V = d.fields.rf_voltage_V.value_or(0.0);
    o.rf_frequency_Hz = d.fields.rf_frequency_Hz.value_or(0.0);
    o.field_array_scale = d.fields.field_array_scale;
    for (int k = 0; k < 3; ++k) o.B_T[k] = d.fields.field_strength_T[k];
    o.temperature_K = e.temperature_K;
    o.pressure_Pa = e.pressure_Pa;
    o.gas_density_m3 = e.particle_density_m_3;
    o.gas_mass_kg = e.gas_mass_kg;
    o.gas_radius_m = e.gas_radius_m;
    o.gas_polarizability_m3 = e.gas_polarizability_m3;
    const int flow = flow_kind(e.flow_model);
    if (flow == 2 && !gas_flow_applied)
        throw Unsupported("domain '" + d.name + "': the axial parabolic flow model needs icarion_gpu_set_gas_flow "
                          "(to_desc with gas_flow_applied)");
    if (flow == 2) {
        for (int k = 0; k < 3; ++k) o.gas_velocity_m_s[k] = 0.0;
    } else if (flow == 1) {
        // gas_velocity_at's AxialUniform branch: to_global_vel((0, 0, u)), the
        // local-to-global rotation's third column times u (the collision
        // handlers on the engine's CPU path and DampingForce both use it)
        const auto& R = d.geometry.rotation_local_to_global;
        for (int k = 0; k < 3; ++k) o.gas_velocity_m_s[k] = R[3 * k + 2] * e.axial_flow_velocity_m_s;
    } else {
        for (int k = 0; k < 3; ++k) o.gas_velocity_m_s[k] = e.gas_velocity_m_s[k];
    }
    o.detector_z_m = detector_z_m;
    return o;
}

// precision: enum icarion_precision; ehss_mode: enum icarion_ehss_mode.
// Collision stages per engine step of the engine's CPU path: subcycles, or in
// multi-event mode max(subcycles, max events) (SimulationEngineCollisionRuntime.cpp:26-33).
inline int32_t collision_substeps(const Physics& p) {
    const int configured = std::max(1, p.collision_subcycles_per_step);
    const int min_multi = std::max(1, p.collision_max_events_per_step);
    return p.collision_multi_event_mode ? std::max(configured, min_multi) : configured;
}

inline icarion_physics_desc to_physics(const Simulation& s, const Physics& p,
                                       int32_t precision = ICARION_PRECISION_FP64,
                                       int32_t ehss_mode = ICARION_EHSS_GEOMETRY) {
    icarion_physics_desc o{};
    o.integrator = integrator_code(s.integrator);
    o.collision_model = collision_code(p.collision_model);
    o.ou_thermal_kicks = p.ou_thermal_kicks ? 1 : 0;
    o.precision = precision;
    o.ehss_mode = ehss_mode;
    o.collision_substeps = collision_substeps(p);
    o.rng_seed = s.rng_seed;
    return o;
}

// ---- species (SpeciesConfig + Molecule) ---------------------------------------
struct Atom {                           // io::Molecule atom, SI units
    std::array<double, 3> pos_m{{0.0, 0.0, 0.0}};
    double mass_u = 0.0;
    double LJ_sigma_m = 0.0;
};

struct SpeciesRecord {
    std::string id;
    bool in_database = true;            // found in the species database
    double mass_amu = 0.0;
    int charge = 1;
    std::optional<double> CCS_A2;       // ion CCS (reference gas)
    std::optional<double> mobility_cm2Vs;
    std::optional<std::string> ccs_reference_gas;
    std::map<std::string, double> ccs_hss_m2;   // CCS_HSS per gas (A^2 in the file, m^2 here)
    std::map<std::string, double> ccs_ehss_m2;  // CCS_EHSS per gas
    std::vector<Atom> atoms;            // geometry_file's molecule (empty: none)

    double CCS_m2() const { return CCS_A2 ? *CCS_A2 * kAngstrom2ToM2 : 0.0; }
};

// sigma for a target gas from a reference-gas sigma: r_ion = max(0,
// sqrt(sigma_ref/pi) - r_ref), sigma = pi (r_ion + r_target)^2; 0 when either
// gas is not tabulated.
inline double derive_ccs_for_target_gas(double sigma_ref_m2, const std::string& gas_ref,
                                        const std::string& gas_target) {
    GasProps r, t;
    if (!tabulated_gas(gas_ref, r) || !tabulated_gas(gas_target, t)) return 0.0;
    const double r_ion = std::max(0.0, std::sqrt(sigma_ref_m2 / kPi) - r.radius_m);
    return kPi * (r_ion + t.radius_m) * (r_ion + t.radius_m);
}

inline double db_value(const std::map<std::string, double>& m, const std::string& gas) {
    const auto it = m.find(gas);
    return it != m.end() && it->second > 0.0 ? it->second : 0.0;
}

// HSS single-gas sigma: CCS_HSS[gas] > derived from the reference gas (when it
// differs from the gas) > the ion's CCS.  0 = no collisions for this species
// (the handler returns false for sigma <= 0).
inline double hss_sigma(const SpeciesRecord& s, const std::string& gas) {
    double sigma = s.CCS_m2();
    if (!s.in_database) return sigma;
    const double db = db_value(s.ccs_hss_m2, gas);
    if (db > 0.0) return db;
    if (s.CCS_m2() > 0.0 && s.ccs_reference_gas && *s.ccs_reference_gas != gas) {
        const double derived = derive_ccs_for_target_gas(s.CCS_m2(), *s.ccs_reference_gas, gas);
        if (derived > 0.0) sigma = derived;
    }
    return sigma;
}

// HSS sigma of one mixture component: CCS_HSS[component] > derived from the
// reference gas > the component's cross_section_m2 > the ion's CCS; throws
// where the handler throws (no sigma for a participating component).
inline double hss_mixture_sigma(const SpeciesRecord& s, const GasComponent& c) {
    double sigma = c.cross_section_m2 > 0.0 ? c.cross_section_m2 : s.CCS_m2();
    if (s.in_database) {
        const double db = db_value(s.ccs_hss_m2, c.species);
        if (db > 0.0) {
            sigma = db;
        } else if (s.CCS_m2() > 0.0 && s.ccs_reference_gas) {
            const double derived = derive_ccs_for_target_gas(s.CCS_m2(), *s.ccs_reference_gas, c.species);
            if (derived > 0.0) sigma = derived;
        }
    }
    if (!(sigma > 0.0))
        throw std::invalid_argument("[HSS] missing CCS for ion '" + s.id + "' and gas '" + c.species + "'");
    return sigma;
}

// Shoemake random rotation from three uniforms, row-major
// (CollisionGeometry::generate_random_rotation, CollisionGeometry.cpp:30-59).
inline void random_rotation(double u1, double u2, double u3, double R[9]) {
    const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    const double x = a * std::sin(2.0 * kPi * u2), y = a * std::cos(2.0 * kPi * u2);
    const double z = b * std::sin(2.0 * kPi * u3), w = b * std::cos(2.0 * kPi * u3);
    R[0] = 1 - 2 * (y * y + z * z);
    R[1] = 2 * (x * y - z * w);
    R[2] = 2 * (x * z + y * w);
    R[3] = 2 * (x * y + z * w);
    R[4] = 1 - 2 * (x * x + z * z);
    R[5] = 2 * (y * z - x * w);
    R[6] = 2 * (x * z - y * w);
    R[7] = 2 * (y * z + x * w);
    R[8] = 1 - 2 * (x * x + y * y);
}

// The hard-sphere atoms of a molecule: centres relative to the mass-weighted
// centre (or the file origin when the masses sum to 0), radius = LJ sigma / 2.
struct Sphere {
    double x, y, z, r;
};

inline std::vector<Sphere> hard_spheres(const std::vector<Atom>& atoms, bool center_on_com = true) {
    double cx = 0.0, cy = 0.0, cz = 0.0;
    if (center_on_com && !atoms.empty()) {
        double m = 0.0;
        for (const Atom& a : atoms) {
            cx += a.pos_m[0] * a.mass_u;
            cy += a.pos_m[1] * a.mass_u;
            cz += a.pos_m[2] * a.mass_u;
            m += a.mass_u;
        }
        if (m > 0.0) {
            cx /= m;
            cy /= m;
            cz /= m;
        } else {
            cx = cy = cz = 0.0;
        }
    }
    std::vector<Sphere> out;
    out.reserve(atoms.size());
    for (const Atom& a : atoms)
        out.push_back(Sphere{a.pos_m[0] - cx, a.pos_m[1] - cy, a.pos_m[2] - cz, 0.5 * a.LJ_sigma_m});
    return out;
}

// Orientation-averaged projection approximation over 64 orientations drawn
// from std::mt19937_64(12345) and uniform_real_distribution<double>(0, 1)
// (the handler's PhysicsRng(12345)): mean of pi Rmax^2, Rmax = max over atoms of
// the rotated centre's distance from the z axis + r_atom + r_gas
// (EHSSCollisionHandler.cpp:474-498).
inline double oapa_sigma(const std::vector<Sphere>& atoms, double neutral_radius_m) {
    std::mt19937_64 gen(12345);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    double A_sum = 0.0;
    for (int k = 0; k < 64; ++k) {
        const double u1 = u01(gen), u2 = u01(gen), u3 = u01(gen);
        double R[9];
        random_rotation(u1, u2, u3, R);
        double Rmax = 0.0;
        for (const Sphere& s : atoms) {
            const double cx = R[0] * s.x + R[1] * s.y + R[2] * s.z;
            const double cy = R[3] * s.x + R[4] * s.y + R[5] * s.z;
            Rmax = std::max(Rmax, std::sqrt(cx * cx + cy * cy) + s.r + neutral_radius_m);
        }
        A_sum += kPi * Rmax * Rmax;
    }
    return A_sum / 64.0;
}

// EHSS sigma_eff for a gas: CCS_EHSS[gas] > OAPA of the geometry > derived from
// the reference gas > the ion's CCS; throws where the handler throws.
inline double ehss_sigma(const SpeciesRecord& s, const std::string& gas, double neutral_radius_m) {
    if (s.in_database) {
        const double db = db_value(s.ccs_ehss_m2, gas);
        if (db > 0.0) return db;
    }
    if (!s.atoms.empty()) return oapa_sigma(hard_spheres(s.atoms), neutral_radius_m);
    if (s.in_database && !gas.empty() && s.CCS_m2() > 0.0 && s.ccs_reference_gas) {
        const double derived = derive_ccs_for_target_gas(s.CCS_m2(), *s.ccs_reference_gas, gas);
        if (derived > 0.0) return derived;
    }
    if (s.CCS_m2() > 0.0) return s.CCS_m2();
    throw std::invalid_argument("[EHSS] no CCS data for species '" + s.id + "' and gas '" + gas + "'");
}

// Species table of one domain and its EHSS atom table (float x, y, z, r per
// atom, the device's atoms_dtype 0).  Species index i of the table is the
// ensemble's species id i.
struct SpeciesTable {
    std::vector<icarion_species_desc> species;
    std::vector<float> atoms_xyzr;
    // mixture cross sections, row-major [species][component], the components
    // and their contact radii (EHSS) (all empty without a mixture):
    // icarion_gpu_set_gas_mixture / icarion_gpu_set_gas_radii
    std::vector<double> mixture_sigma_m2;
    std::vector<icarion_gas_component> mixture;
    std::vector<double> mixture_radius_m;

    uint32_t n_atoms() const { return static_cast<uint32_t>(atoms_xyzr.size() / 4); }
};

inline SpeciesTable species_table(const std::vector<SpeciesRecord>& records, const Domain& d,
                                  const Physics& p) {
    const int32_t model = collision_code(p.collision_model);
    const Environment& e = d.env;
    if ((model == ICARION_COLLISION_LANGEVIN || model == ICARION_COLLISION_FRICTION) && !e.gas_mixture.empty())
        throw Unsupported("damping over a gas mixture (Blanc's law) has no device path");
    SpeciesTable t;
    for (const SpeciesRecord& s : records) {
        icarion_species_desc o{};
        o.ccs_m2 = s.CCS_m2();
        o.K0_cm2_Vs = s.mobility_cm2Vs.value_or(0.0);
        if (model == ICARION_COLLISION_HSS && e.gas_mixture.empty()) o.sigma_hss_m2 = hss_sigma(s, e.gas_species);
        if (model == ICARION_COLLISION_EHSS) {
            // single gas: the domain gas and radius; a mixture carries its own
            // per-component values below
            if (e.gas_mixture.empty()) o.sigma_ehss_m2 = ehss_sigma(s, e.gas_species, e.gas_radius_m);
            if (!s.atoms.empty()) {
                o.atom_offset = t.n_atoms();
                o.atom_count = static_cast<uint32_t>(s.atoms.size());
                for (const Sphere& a : hard_spheres(s.atoms)) {
                    t.atoms_xyzr.push_back(static_cast<float>(a.x));
                    t.atoms_xyzr.push_back(static_cast<float>(a.y));
                    t.atoms_xyzr.push_back(static_cast<float>(a.z));
                    t.atoms_xyzr.push_back(static_cast<float>(a.r));
                }
            }
        }
        t.species.push_back(o);
    }
    if ((model == ICARION_COLLISION_HSS || model == ICARION_COLLISION_EHSS) && !e.gas_mixture.empty()) {
        if (e.gas_mixture.size() > ICARION_MAX_GAS_COMPONENTS)
            throw Unsupported("more than " + std::to_string(ICARION_MAX_GAS_COMPONENTS) + " gas components");
        for (const GasComponent& c : e.gas_mixture) {
            t.mixture.push_back(icarion_gas_component{c.density_m3, c.mass_kg, c.participates_in_collisions ? 1 : 0, 0});
            t.mixture_radius_m.push_back(c.radius_m > 0.0 ? c.radius_m : e.gas_radius_m);
        }
        for (const SpeciesRecord& s : records) {
            for (size_t k = 0; k < e.gas_mixture.size(); ++k) {
                const GasComponent& c = e.gas_mixture[k];
                double sig = 0.0;
                if (model == ICARION_COLLISION_HSS) {
                    // the HSS handler skips components that do not take part or have no density
                    if (c.participates_in_collisions && c.density_m3 > 0.0) sig = hss_mixture_sigma(s, c);
                } else if (c.participates_in_collisions) {
                    // the EHSS handler's mixture path: compute_effective_ccs with the
                    // component's radius (else the domain's) for every participating one
                    sig = ehss_sigma(s, c.species, t.mixture_radius_m[k]);
                }
                t.mixture_sigma_m2.push_back(sig);
            }
        }
    } else {
        t.mixture_radius_m.clear();
    }
    return t;
}

// ---- the fused step's eligibility --------------------------------------------
// The device step runs push, collisions and boundary elimination of every ion
// in one call, the collision stage as often per step as the engine's CPU path
// would (collision_substeps: subcycles, multi-event mode).  The engine takes
// its batch collision path only when no multi-event mode, one subcycle, no
// deep collision diagnostics, a constant flow model in every domain and one
// dt for all active ions hold; otherwise it collides ions on the CPU
// (handle_collisions_cpu), which after a fused step would collide them twice.
// So a reference-side collision handler skips its per-ion handle_collision for
// the ions of a fused step (they have had their substeps), and the fused step
// itself is eligible when the device covers what the CPU path would do: no
// deep collision diagnostics (the per-event records are host-side), constant
// or axial uniform flow in every domain, one dt for all active ions.  A reference-side strategy
// answers supports_batch() / step_batch() with this predicate and leaves the
// whole step to the CPU when it is false.
struct EngineConditions {
    bool collision_multi_event_mode = false;
    int collision_subcycles_per_step = 1;
    int collision_max_events_per_step = 1;
    bool deep_collision_diagnostics = false;
    bool all_flow_models_constant = true;   // EnvironmentConfig::flow_model == Constant everywhere
    bool all_flow_models_covered = true;    // flow_covered() in every domain
    bool uniform_dt = true;                 // every active ion has the same dt
};

// the engine's own batch-collision preconditions (SimulationEngineCollisionRuntime.cpp:46)
inline bool engine_batch_collisions(const EngineConditions& c) {
    const int configured = std::max(1, c.collision_subcycles_per_step);
    const int min_multi = std::max(1, c.collision_max_events_per_step);
    const int subcycles = c.collision_multi_event_mode ? std::max(configured, min_multi) : configured;
    return !c.collision_multi_event_mode && subcycles == 1 && !c.deep_collision_diagnostics &&
           c.all_flow_models_constant && c.uniform_dt;
}

inline bool fused_step_eligible(const EngineConditions& c) {
    return !c.deep_collision_diagnostics && c.all_flow_models_covered && c.uniform_dt;
}

// all_flow_models_covered: flow_covered() in every domain (by default the
// same as all_flow_models_constant).
inline EngineConditions engine_conditions(const Physics& p, bool deep_collision_diagnostics,
                                          bool all_flow_models_constant, bool uniform_dt,
                                          std::optional<bool> all_flow_models_covered = std::nullopt) {
    EngineConditions c;
    c.collision_multi_event_mode = p.collision_multi_event_mode;
    c.collision_subcycles_per_step = p.collision_subcycles_per_step;
    c.collision_max_events_per_step = p.collision_max_events_per_step;
    c.deep_collision_diagnostics = deep_collision_diagnostics;
    c.all_flow_models_constant = all_flow_models_constant;
    c.all_flow_models_covered = all_flow_models_covered.value_or(all_flow_models_constant);
    c.uniform_dt = uniform_dt;
    return c;
}

}  // namespace config
}  // namespace icarion

#endif  // ICARION_CONFIG_HPP
