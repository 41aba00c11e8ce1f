// config_driver.cpp - fills the plain config structs of include/icarion_config.hpp
// from a reference configuration file and its species database (field names of
// the reference's loaders: SimulationConfigLoader, PhysicsConfigLoader,
// DomainConfigLoader, SpeciesLoader, moleculeLoader), runs the helpers and prints
// the resulting descriptors as one JSON document for tests/test_config.py.
//
// usage: config_driver CONFIG SPECIES_DB MOLECULE_DIR [COLLISION_MODEL]
#include <cstdio>
#include <iostream>
#include <string>

#include "../../include/icarion_config.hpp"
#include "mini_json.hpp"

namespace C = icarion::config;
using mini_json::Value;

static std::array<double, 3> vec3(const Value& v, std::array<double, 3> dflt = {{0, 0, 0}}) {
    if (v.kind != Value::Array || v.size() != 3) return dflt;
    return {{v[0].as_double(), v[1].as_double(), v[2].as_double()}};
}

static std::optional<double> opt(const Value& v) {
    if (v.is_number()) return v.as_double();
    return std::nullopt;
}

static C::Domain load_domain(const Value& j) {
    C::Domain d;
    d.name = j["name"].as_string();
    d.instrument = j["instrument"].as_string("IMS");
    const Value& g = j["geometry"];
    d.geometry.origin_m = vec3(g["origin_m"]);
    d.geometry.length_m = g["length_m"].as_double();
    d.geometry.radius_m = g["radius_m"].as_double();
    d.geometry.radius_in_m = g["radius_in_m"].as_double();
    d.geometry.radius_out_m = g["radius_out_m"].as_double();
    d.geometry.radius_char_m = g["radius_char_m"].as_double();
    d.geometry.end_aperture_m = g["end_aperture_m"].as_double();
    const Value& e = j["env"];
    d.env.temperature_K = e["temperature_K"].as_double(300.0);
    d.env.pressure_Pa = e["pressure_Pa"].as_double(101325.0);
    d.env.gas_species = e["gas_species"].as_string("He");
    d.env.gas_velocity_m_s = vec3(e["gas_velocity_m_s"]);
    // flow_model and flow_parameters as DomainConfigLoader::load_environment reads them
    // (DomainConfigLoader.cpp:204-221)
    d.env.flow_model = e["flow_model"].as_string("constant");
    d.env.axial_flow_velocity_m_s = e["axial_flow_velocity_m_s"].as_double(0.0);
    d.env.axial_flow_max_velocity_m_s = e["axial_flow_max_velocity_m_s"].as_double(0.0);
    {
        const Value& fp = e["flow_parameters"];
        d.env.axial_flow_velocity_m_s = fp["axial_flow_velocity_m_s"].as_double(d.env.axial_flow_velocity_m_s);
        d.env.axial_flow_max_velocity_m_s =
            fp["axial_flow_max_velocity_m_s"].as_double(d.env.axial_flow_max_velocity_m_s);
    }
    const Value& mix = e["gas_mixture"];
    for (size_t k = 0; k < mix.size(); ++k) {
        C::GasComponent c;
        c.species = mix[k]["species"].as_string();
        c.mole_fraction = mix[k]["mole_fraction"].as_double();
        c.cross_section_m2 = mix[k]["cross_section_m2"].as_double(-1.0);
        c.polarizability_m3 = mix[k]["polarizability_m3"].as_double(-1.0);
        c.participates_in_collisions = mix[k]["participates_in_collisions"].as_bool(true);
        d.env.gas_mixture.push_back(c);
    }
    d.env.derive();
    const Value& f = j["fields"];
    d.fields.EN_Td = opt(f["DC"]["EN_Td"]);
    d.fields.axial_V = opt(f["DC"]["axial_V"]);
    d.fields.quad_V = opt(f["DC"]["quad_V"]);
    d.fields.radial_V = opt(f["DC"]["radial_V"]);
    d.fields.rf_voltage_V = opt(f["RF"]["voltage_V"]);
    d.fields.rf_frequency_Hz = opt(f["RF"]["frequency_Hz"]);
    d.fields.rf_phase_rad = f["RF"]["phase_rad"].as_double();
    const Value& m = f.has("B") ? f["B"] : f["magnetic"];
    d.fields.magnetic_enabled = m["enabled"].as_bool(false) || m.has("field_strength_T");
    d.fields.field_strength_T = vec3(m["field_strength_T"]);
    d.fields.field_array = f.has("field_array_terms");
    return d;
}

static C::SpeciesRecord load_species(const std::string& id, const Value& db, const std::string& db_dir) {
    C::SpeciesRecord s;
    s.id = id;
    const Value& e = db["species"][id];
    s.in_database = e.kind == Value::Object;
    if (!s.in_database) return s;
    s.mass_amu = e["mass_amu"].as_double();
    s.charge = static_cast<int>(e["charge"].as_double(1.0));
    s.CCS_A2 = opt(e["CCS_A2"]);
    s.mobility_cm2Vs = opt(e["mobility_cm2Vs"]);
    if (e["CCS_reference_gas"].kind == Value::String) s.ccs_reference_gas = e["CCS_reference_gas"].as_string();
    for (const auto& kv : e["CCS_HSS"].obj) s.ccs_hss_m2[kv.first] = kv.second.as_double() * C::kAngstrom2ToM2;
    for (const auto& kv : e["CCS_EHSS"].obj) s.ccs_ehss_m2[kv.first] = kv.second.as_double() * C::kAngstrom2ToM2;
    if (e["geometry_file"].kind == Value::String) {
        std::string path = db_dir + "/" + e["geometry_file"].as_string();
        const Value mol = mini_json::parse_file(path)["molecule"];
        for (size_t k = 0; k < mol["atoms"].size(); ++k) {
            const Value& a = mol["atoms"][k];
            C::Atom at;
            const auto p = vec3(a["pos"]);
            for (int c = 0; c < 3; ++c) at.pos_m[c] = p[c] * C::kAngstromToM;
            at.mass_u = a["mass_u"].as_double();
            at.LJ_sigma_m = a["LJ_sigma_angstrom"].as_double() * C::kAngstromToM;
            s.atoms.push_back(at);
        }
    }
    return s;
}

static void num(std::ostream& o, const char* k, double v, bool comma = true) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    o << '"' << k << "\": " << (std::isnan(v) ? std::string("NaN") : std::string(buf)) << (comma ? ", " : "");
}

int main(int argc, char** argv) {
    if (argc == 2 && std::string(argv[1]) == "--gate") {
        // fused_step_eligible, the engine's own batch-collision predicate and
        // the collision substeps over every combination of the engine conditions
        std::cout << "[";
        bool first = true;
        for (int multi = 0; multi < 2; ++multi)
            for (int sub = 1; sub <= 2; ++sub)
                for (int maxev = 1; maxev <= 3; maxev += 2)
                    for (int deep = 0; deep < 2; ++deep)
                        for (int flow = 0; flow < 2; ++flow)
                            for (int udt = 0; udt < 2; ++udt) {
                                C::Physics ph;
                                ph.collision_multi_event_mode = multi != 0;
                                ph.collision_subcycles_per_step = sub;
                                ph.collision_max_events_per_step = maxev;
                                const C::EngineConditions ec =
                                    C::engine_conditions(ph, deep != 0, flow != 0, udt != 0);
                                const bool ok = C::fused_step_eligible(ec);
                                const bool engine = C::engine_batch_collisions(ec);
                                std::cout << (first ? "" : ", ") << "[" << multi << ", " << sub << ", " << maxev
                                          << ", " << deep << ", " << flow << ", " << udt << ", " << (ok ? 1 : 0)
                                          << ", " << (engine ? 1 : 0) << ", " << C::collision_substeps(ph) << "]";
                                first = false;
                            }
        std::cout << "]\n";
        return 0;
    }
    if (argc == 2 && std::string(argv[1]) == "--flow") {
        // to_desc's gas velocity for the flow models: the rotated AxialUniform case
        // of the reference's engine test (test_simulation_engine.cpp:652-685:
        // u = 100 m/s, local-to-global columns (0,0,-1), (0,1,0), (1,0,0)), an
        // unrotated one, the constant model, and whether AxialParabolic is taken
        C::Domain d;
        d.name = "flow";
        d.geometry.length_m = 0.02;
        d.geometry.radius_m = 0.01;
        d.env.gas_species = "N2";
        d.env.derive();
        d.fields.axial_V = 10.0;
        auto show = [&](const char* tag) {
            std::cout << "\"" << tag << "\": ";
            try {
                const icarion_domain_desc o = C::to_desc(d);
                std::cout << "[" << o.gas_velocity_m_s[0] << ", " << o.gas_velocity_m_s[1] << ", "
                          << o.gas_velocity_m_s[2] << "]";
            } catch (const C::Unsupported&) {
                std::cout << "\"unsupported\"";
            }
        };
        std::cout << "{";
        d.env.flow_model = "axial_uniform";
        d.env.axial_flow_velocity_m_s = 100.0;
        d.env.gas_velocity_m_s = {{0.0, 0.0, 100.0}};
        d.geometry.rotation_local_to_global = {{0, 0, 1, 0, 1, 0, -1, 0, 0}};   // row-major
        d.geometry.rotation_global_to_local = {{0, 0, -1, 0, 1, 0, 1, 0, 0}};
        show("uniform_rotated");
        std::cout << ", ";
        d.geometry.rotation_local_to_global = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
        d.geometry.rotation_global_to_local = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
        d.env.flow_model = "Uniform";
        d.env.axial_flow_velocity_m_s = 134.0;
        show("uniform");
        std::cout << ", ";
        d.env.flow_model = "legacy";
        d.env.gas_velocity_m_s = {{1.0, 2.0, 3.0}};
        show("constant");
        std::cout << ", ";
        d.env.flow_model = "poiseuille";
        d.env.axial_flow_max_velocity_m_s = 250.0;
        show("parabolic");
        const icarion_domain_desc pa = C::to_desc(d, NAN, true);
        const C::GasFlow gf = C::gas_flow(d.env);
        std::cout << ", \"parabolic_applied\": [" << pa.gas_velocity_m_s[0] << ", " << pa.gas_velocity_m_s[1]
                  << ", " << pa.gas_velocity_m_s[2] << "], \"parabolic_flow\": [" << gf.model << ", "
                  << gf.axial_max_velocity_m_s << "]";
        std::cout << ", \"covered\": [" << C::flow_covered(d.env) << ", " << C::flow_covered(d.env, false);
        d.env.flow_model = "axial_const";
        const C::GasFlow gu = C::gas_flow(d.env);
        std::cout << ", " << C::flow_covered(d.env) << ", " << C::flow_covered(d.env, false) << "], \"uniform_flow\": ["
                  << gu.model << ", " << gu.axial_max_velocity_m_s << "]}\n";
        return 0;
    }
    if (argc < 4) {
        std::cerr << "usage: config_driver CONFIG SPECIES_DB MOLECULE_DIR [COLLISION_MODEL]\n";
        return 2;
    }
    try {
        const Value cfg = mini_json::parse_file(argv[1]);
        const Value db = mini_json::parse_file(argv[2]);
        const std::string db_dir = argv[3];
        C::Simulation sim;
        sim.integrator = cfg["simulation"]["integrator"].as_string("RK4");
        sim.dt_s = cfg["simulation"]["dt_s"].as_double();
        sim.rng_seed = static_cast<uint64_t>(cfg["simulation"]["rng_seed"].as_double());
        sim.enable_gpu = cfg["simulation"]["enable_gpu"].as_bool(false);
        C::Physics phys;
        phys.collision_model = argc > 4 ? argv[4] : cfg["physics"]["collision_model"].as_string();
        phys.collision_multi_event_mode = cfg["physics"]["collision_multi_event_mode"].as_bool(false);
        phys.collision_subcycles_per_step =
            static_cast<int>(cfg["physics"]["collision_subcycles_per_step"].as_double(1.0));
        const icarion_physics_desc p = C::to_physics(sim, phys);
        std::vector<C::SpeciesRecord> recs;
        const Value& ions = cfg["ions"]["species"];
        for (size_t k = 0; k < ions.size(); ++k) recs.push_back(load_species(ions[k]["id"].as_string(), db, db_dir));
        std::ostream& o = std::cout;
        o << "{\"physics\": {";
        o << "\"integrator\": " << p.integrator << ", \"collision_model\": " << p.collision_model
          << ", \"rng_seed\": " << p.rng_seed << "}, \"domains\": [";
        const Value& doms = cfg["domains"];
        for (size_t k = 0; k < doms.size(); ++k) {
            const C::Domain d = load_domain(doms[k]);
            const icarion_domain_desc x = C::to_desc(d);
            o << (k ? ", " : "") << "{\"name\": \"" << d.name << "\", \"desc\": {";
            o << "\"instrument\": " << x.instrument << ", \"use_field_array\": " << x.use_field_array
              << ", \"magnetic_enabled\": " << x.magnetic_enabled << ", ";
            auto arr = [&](const char* key, const double* v, int n) {
                o << '"' << key << "\": [";
                for (int c = 0; c < n; ++c) {
                    char buf[40];
                    std::snprintf(buf, sizeof(buf), "%.17g", v[c]);
                    o << (c ? ", " : "") << buf;
                }
                o << "], ";
            };
            arr("origin_m", x.origin_m, 3);
            arr("rot_g2l", x.rot_g2l, 9);
            arr("rot_l2g", x.rot_l2g, 9);
            arr("B_T", x.B_T, 3);
            arr("gas_velocity_m_s", x.gas_velocity_m_s, 3);
            num(o, "field_array_scale", x.field_array_scale);
            num(o, "origin_z_m", x.origin_m[2]);
            num(o, "length_m", x.length_m);
            num(o, "radius_m", x.radius_m);
            num(o, "radius_in_m", x.radius_in_m);
            num(o, "radius_out_m", x.radius_out_m);
            num(o, "radius_char_m", x.radius_char_m);
            num(o, "dc_axial_V", x.dc_axial_V);
            num(o, "dc_quad_V", x.dc_quad_V);
            num(o, "dc_radial_V", x.dc_radial_V);
            num(o, "rf_voltage_V", x.rf_voltage_V);
            num(o, "rf_frequency_Hz", x.rf_frequency_Hz);
            num(o, "temperature_K", x.temperature_K);
            num(o, "pressure_Pa", x.pressure_Pa);
            num(o, "gas_density_m3", x.gas_density_m3);
            num(o, "gas_mass_kg", x.gas_mass_kg);
            num(o, "gas_radius_m", x.gas_radius_m);
            num(o, "gas_polarizability_m3", x.gas_polarizability_m3, false);
            o << "}, \"species\": [";
            try {
                const C::SpeciesTable t = C::species_table(recs, d, phys);
                for (size_t s = 0; s < t.species.size(); ++s) {
                    const icarion_species_desc& sp = t.species[s];
                    o << (s ? ", " : "") << "{";
                    num(o, "sigma_hss_m2", sp.sigma_hss_m2);
                    num(o, "sigma_ehss_m2", sp.sigma_ehss_m2);
                    num(o, "ccs_m2", sp.ccs_m2);
                    num(o, "K0_cm2_Vs", sp.K0_cm2_Vs);
                    num(o, "mass_amu", recs[s].mass_amu);
                    o << "\"charge\": " << recs[s].charge << ", ";
                    o << "\"atom_offset\": " << sp.atom_offset << ", \"atom_count\": " << sp.atom_count << "}";
                }
                o << "], \"atoms_xyzr\": [";
                for (size_t a = 0; a < t.atoms_xyzr.size(); ++a) {
                    char buf[32];
                    std::snprintf(buf, sizeof(buf), "%.9g", static_cast<double>(t.atoms_xyzr[a]));
                    o << (a ? ", " : "") << buf;
                }
                o << "], \"mixture\": [";
                for (size_t c = 0; c < t.mixture.size(); ++c) {
                    o << (c ? ", " : "") << "{";
                    num(o, "density_m3", t.mixture[c].density_m3);
                    num(o, "mass_kg", t.mixture[c].mass_kg);
                    o << "\"participates\": " << t.mixture[c].participates << "}";
                }
                o << "], \"mixture_sigma_m2\": [";
                for (size_t c = 0; c < t.mixture_sigma_m2.size(); ++c) {
                    char buf[40];
                    std::snprintf(buf, sizeof(buf), "%.17g", t.mixture_sigma_m2[c]);
                    o << (c ? ", " : "") << buf;
                }
                o << "], \"error\": null}";
            } catch (const C::Unsupported& e) {
                o << "], \"error\": \"unsupported: " << e.what() << "\"}";
            } catch (const std::invalid_argument& e) {
                o << "], \"error\": \"invalid: " << e.what() << "\"}";
            }
        }
        o << "]}\n";
    } catch (const C::Unsupported& e) {
        std::cout << "{\"error\": \"unsupported: " << e.what() << "\"}\n";
        return 3;
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 1;
    }
    return 0;
}
