/*
 * icarion_oracle.h - CPU FP64 restatement of the reference batch ion-step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the
 * host-CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product library (icarion_amd) never links,
 * loads or calls it.
 *
 * Every function restates a reference function step for step; the
 * citation next to each definition in icarion_oracle.cpp names the
 * reference file:line it follows.  Descriptors are the product's own
 * (include/icarion_gpu.h) so both sides read identical inputs.
 */
#ifndef ICARION_ORACLE_H
#define ICARION_ORACLE_H

#include "../include/icarion_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One term of a composite field (CompositeFieldProvider::FieldTerm,
 * CompositeFieldProvider.h:39-46): its components on the field's grid and
 * its time dependence (enum icarion_field_scale). */
typedef struct oracle_field_term {
    int32_t scale_kind;
    int32_t reserved0;
    double constant_scale;
    double frequency_Hz;
    double phase_rad;
    const double *Ex, *Ey, *Ez;
} oracle_field_term;

/* Field array in the reference's own layout (FieldArray,
 * fieldArrayLoader.h:23-40): monotone axes, x-fastest FP64 components.
 * n_terms == 0: the single array (Ex, Ey, Ez), unscaled (GridFieldProvider);
 * n_terms == 1: that term's arrays, unscaled (PhysicsSetup.cpp:165-171);
 * n_terms >= 2: sum_i scale_i(t) E_i (CompositeFieldProvider). */
typedef struct oracle_field {
    uint32_t nx, ny, nz;
    const double *xs, *ys, *zs;
    const double *Ex, *Ey, *Ez;
    uint32_t n_terms;
    uint32_t reserved0;
    const oracle_field_term *terms;
    /* voltage waveforms of the analytic fields per icarion_voltage_slot (NULL:
     * the domain's constant); the grid members are unused when the domain has
     * no field array */
    const icarion_waveform_desc *waveforms[ICARION_N_VOLTAGE_SLOTS];
    /* the domain's gas pressure as a waveform (EnvironmentConfig::
     * pressure_Pa_waveform), NULL: the descriptor's constant pressure */
    const icarion_waveform_desc *pressure;
    /* the domain's gas flow (enum icarion_flow_model) and the AxialParabolic
     * centre-line speed (none: 0) */
    int32_t flow_model;
    int32_t reserved1;
    double flow_max_velocity_m_s;
} oracle_field;

/* ValueOrWaveform::evaluate of one waveform (WaveformConfig.h:33-200,
 * WaveformConfig.cpp:10-91). */
double oracle_waveform_value(const icarion_waveform_desc *w, double t);

/* EnvironmentConfig::gas_velocity_at (EnvironmentConfig.h:96-129) of the
 * domain at the global position pos, with field's flow (NULL: constant). */
void oracle_gas_velocity(const icarion_domain_desc *domain, const oracle_field *field, const double pos[3],
                         double out[3]);

/* DomainManager::find_domain_index (DomainManager.cpp:66-109) over n_domains
 * domains with their end apertures (NULL: the radii). */
int oracle_find_domain(const icarion_domain_desc *domains, const double *end_aperture_m, uint32_t n_domains,
                       const double pos[3]);

/* The engine loop over several domains (analytic fields, constant voltages,
 * RK4 / Boris, Philox streams); domain_index[n] in/out is each ion's domain
 * index (-1 before its first step).  fields: NULL, or per domain NULL or an
 * oracle_field carrying that domain's voltage waveforms.  mixtures: NULL, or
 * per domain NULL or that domain's gas mixture (HSS).  domain_species: NULL,
 * or per domain NULL (the species table) or a table of n_species rows with
 * that domain's cross sections, CCS and K0. */
int oracle_run_multi(const icarion_domain_desc *domains, const double *end_aperture_m, uint32_t n_domains,
                     const oracle_field *const *fields, const struct oracle_gas_mixture *const *mixtures,
                     const icarion_species_desc *const *domain_species,
                     const icarion_physics_desc *physics,
                     const icarion_species_desc *species, uint32_t n_species, const double *atoms,
                     icarion_ion_soa *ions, int32_t *domain_index, double t0, double dt, uint64_t n_steps,
                     uint64_t step_offset, int n_threads, icarion_step_stats *stats);

/* CompositeFieldProvider::compute_scale_factor (CompositeFieldProvider.h:117-171):
 * the domain's constant voltages, or the waveforms of field (may be NULL) by
 * CompositeFieldProvider::eval_voltage's rule (inline waveforms give 0). */
double oracle_field_term_scale(const icarion_domain_desc *domain, const oracle_field_term *term,
                               const oracle_field *field, double t);

/* Gas mixture of the HSS mixture path: components and the per-species
 * cross sections sigma_m2[s * n_components + c]. */
typedef struct oracle_gas_mixture {
    uint32_t n_components;
    uint32_t n_species;
    const icarion_gas_component *components;
    const double *sigma_m2;
    const double *radius_m;   /* per component, EHSS contact radius (NULL / 0: the domain's) */
} oracle_gas_mixture;

/* Offline EHSS sample sets (EHSSOfflineSampleSet, EHSSOfflineSampleSet.h:15-30)
 * per species: species s has n_orient[s] orientations (0: the online geometry
 * method), each with sigma_eff and mu_stride[s] first-contact mu; the tables
 * of all species concatenated in species order. */
typedef struct oracle_ehss_offline {
    uint32_t n_species;
    uint32_t reserved0;
    const uint32_t *n_orient;
    const uint32_t *mu_stride;
    const double *sigma_eff;
    const double *mu;
} oracle_ehss_offline;

/* EHSS orientation samples (EHSSOrientationSamples, EHSSSamples.h:14-17) per
 * species: n_orient[s] orientations as quaternions (w, x, y, z) and their
 * areas for the domain's gas; the tables of all species concatenated. */
typedef struct oracle_ehss_orientations {
    uint32_t n_species;
    uint32_t reserved0;
    const uint32_t *n_orient;
    const double *quat_wxyz;
    const double *area_m2;
} oracle_ehss_orientations;

/* Optional collision tables of a run (any pointer may be NULL). */
typedef struct oracle_collision_tables {
    const oracle_gas_mixture *mixture;
    const oracle_ehss_offline *ehss_offline;
    const oracle_ehss_orientations *ehss_orient;
} oracle_collision_tables;

/* CollisionKernels::ehss_collision_with_orientation (CollisionKernels.cpp:44-137)
 * with the rotation of quaternion_to_rotation (CollisionGeometry.cpp:73-104),
 * sigma_eff and the forced-hit budget, on one MT19937-64 stream seeded with seed
 * (the reference's PhysicsRng). */
void oracle_mt_ehss_oriented(uint64_t seed, const double v_ion[3], const double v_neutral[3], double ion_mass,
                             double neutral_mass, double ion_radius, const double *atoms_xyzr, uint32_t n_atoms,
                             const double quat_wxyz[4], double sigma_eff, double out[3]);
void oracle_quaternion_to_rotation(const double quat_wxyz[4], double R[9]);

/* One EHSS collision stage over a gas mixture (EHSSCollisionHandler.cpp:311-424,
 * geometry method) for an ion of the given species, on one MT19937-64 stream;
 * vel in/out, returns 1 on a collision. */
int oracle_mt_ehss_mixture(uint64_t seed, const icarion_domain_desc *domain, const oracle_gas_mixture *mix,
                           uint32_t species, const double *atoms_xyzr, uint32_t n_atoms, double mass_kg,
                           double vel[3], double dt);

enum { ORACLE_RNG_PHILOX = 0, ORACLE_RNG_MT19937_64 = 1 };

/* Philox4x32-10 block function (Salmon et al., SC'11). */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* The FP64 uniform the Philox stream hands out for words (hi, lo). */
double oracle_uniform_from_words(uint32_t hi, uint32_t lo);

/* First n FP64 uniforms of the (seed, ion, step) Philox stream. */
void oracle_philox_uniforms(uint64_t seed, uint32_t ion, uint64_t step, uint32_t n, double *out);

/* E(pos, t) of the domain (analytic model, or field array when
 * domain->use_field_array and field != NULL). */
void oracle_field_E(const icarion_domain_desc *domain, const oracle_field *field,
                    const double pos[3], double t, double E_out[3]);

/* Domain containment (CylindricalGeometry::contains / OrbitrapGeometry::contains). */
int oracle_contains(const icarion_domain_desc *domain, const double pos[3]);

/* Orbitrap electrode radius at axial position z by bisection
 * (OrbitrapGeometry.h:181-219). */
double oracle_orbitrap_r_for_z(double z, double R, double R_m);

/* Trilinear interpolation with clamping (fieldArrayLoader.cpp:103-162). */
void oracle_interpolate_field(const oracle_field *field, const double pos[3], double E_out[3]);

/* Damping coefficient gamma of DampingForce for the collision model
 * (Langevin / Friction), DampingForce.cpp:105-277. */
double oracle_damping_gamma(const icarion_domain_desc *domain, int collision_model,
                            const icarion_species_desc *sp, double mass_kg,
                            double charge_C, const double vel[3]);

/* ---- single kernels driven by std::mt19937_64 (PhysicsRng(seed)) ------
 * These consume random numbers in exactly the reference order so they can be
 * compared bit for bit with the reference's compiled kernels. */
void oracle_mt_sample_neutral_velocity(uint64_t seed, double T_K, double mass_kg,
                                       const double flow[3], double out[3]);
void oracle_mt_hss_collision(uint64_t seed, const double v_ion[3], const double v_neutral[3],
                             double ion_mass, double neutral_mass, double out[3]);
void oracle_mt_ehss_collision(uint64_t seed, const double v_ion[3], const double v_neutral[3],
                              double ion_mass, double neutral_mass, double ion_radius,
                              const double *atoms_xyzr, uint32_t n_atoms, double out[3]);
/* OAPA sigma_eff over 64 orientations from PhysicsRng(12345)
 * (EHSSCollisionHandler.cpp:476-497). */
double oracle_oapa_sigma(const double *atoms_xyzr, uint32_t n_atoms, double neutral_radius);

/* One RK4 step / Boris step of a single ion (no collisions, no boundary),
 * for unit tests of the push against closed forms. */
void oracle_push_one(const icarion_domain_desc *domain, const icarion_physics_desc *physics,
                     const icarion_species_desc *sp, const oracle_field *field,
                     double mass_kg, double charge_C, double pos[3], double vel[3],
                     double t, double dt);

/* The whole engine loop: n_steps steps of dt from engine time t0 over the
 * ensemble in place.  rng_kind selects Philox streams (identical to the
 * device) or per-ion std::mt19937_64 seeded rng_seed + i (RngUtils.cpp:51-53).
 * n_threads <= 0 uses the OpenMP default.  Returns 0 on success. */
int oracle_run(const icarion_domain_desc *domain, const icarion_physics_desc *physics,
               const icarion_species_desc *species, uint32_t n_species,
               const double *atoms_xyzr, uint32_t n_atoms, const oracle_field *field,
               icarion_ion_soa *ions, double t0, double dt, uint64_t n_steps,
               uint64_t step_offset, int rng_kind, int n_threads,
               icarion_step_stats *stats, const oracle_collision_tables *tables);

/* ---- RK45 (RK45Strategy.cpp:101-520) -------------------------------------
 * Per-ion adaptive state carried between engine steps (RK45Strategy::RK45State,
 * RK45Strategy.h:192-198 + the engine's dt_per_ion_). */
typedef struct oracle_rk45_ion {
    double t;            /* ion time t_i                                   */
    double dt;           /* next trial dt (dt_per_ion_)                     */
    double last_error;   /* PI memory (1.0 initially)                       */
    double k1a[3];       /* FSAL acceleration of the last accepted step     */
    int32_t fsal;        /* k1a valid                                       */
    int32_t reserved;
} oracle_rk45_ion;

/* err = max over x, y, z, vx, vy, vz of |y5 - y4| / (atol + rtol |y0|),
 * floored at 1e-10 (RK45Strategy::estimate_error, :272-313). */
double oracle_rk45_error(const icarion_rk45_desc *cfg, const double y4[6], const double y5[6],
                         const double y0[6]);

/* PI step-size controller (RK45Strategy::compute_new_step, :315-351). */
double oracle_rk45_new_step(const icarion_rk45_desc *cfg, double dt, double err, double dt_min,
                            double dt_max, double prev_err);

/* One RK45Strategy::step of one ion from engine time t with trial dt (no
 * collisions, no boundary): pos/vel updated in place, st updated (FSAL,
 * last_error); returns the accepted dt, *dt_next the suggestion, *rejected the
 * rejected trials; -1 if no trial was accepted (the reference throws). */
double oracle_rk45_push_one(const icarion_domain_desc *domain, const icarion_physics_desc *physics,
                            const icarion_rk45_desc *cfg, const icarion_species_desc *sp,
                            const oracle_field *field, double mass_kg, double charge_C,
                            double pos[3], double vel[3], double t, double dt,
                            oracle_rk45_ion *st, double *dt_next, int *rejected);

/* The engine loop with RK45: n_steps engine steps from engine time t0.  Each
 * step integrates every live ion from the engine time with its own dt_i, uses
 * the accepted dt for its collision and its time t_i, and then takes the
 * engine time as max_i t_i (SimulationEngine.cpp:374-428,
 * SimulationEngineIntegration.cpp:110-146, SimulationEngineLifecycle.cpp:172,
 * SimulationEngineUtils.cpp:49-83).  rk45[] holds the per-ion state in/out
 * (its t, dt replace ions->time_s / dt_s).  Random numbers: the Philox streams
 * of the device.  stats->t_end_s = the engine time after the last step. */
int oracle_run_rk45(const icarion_domain_desc *domain, const icarion_physics_desc *physics,
                    const icarion_rk45_desc *cfg, const icarion_species_desc *species,
                    uint32_t n_species, const double *atoms_xyzr, uint32_t n_atoms,
                    const oracle_field *field, icarion_ion_soa *ions, oracle_rk45_ion *rk45,
                    double t0, double dt, uint64_t n_steps, uint64_t step_offset, int n_threads,
                    icarion_step_stats *stats, const oracle_collision_tables *tables);

/* One mixture collision stage (HSSCollisionHandler.cpp:105-249) of an ion of
 * mass m with velocity vel (in/out) on a PhysicsRng seeded with seed
 * (mt19937_64).  sigma_row: the ion's cross sections per component.  Returns
 * the selected component (or -1 without a collision); vn_out (may be NULL)
 * receives the selected neutral's velocity. */
int oracle_mt_hss_mixture(uint64_t seed, const icarion_domain_desc *domain,
                          const oracle_gas_mixture *mixture, const double *sigma_row, double mass_kg,
                          double vel[3], double dt, double vn_out[3]);

/* One EHSS collision stage with an offline sample set (EHSSCollisionHandler.cpp:
 * 151-172, 222-299: orientation drawn first, sigma_eff of that orientation in
 * the collision test, mu from its samples, contact normal at a uniform
 * azimuth, specular reflection) on a PhysicsRng seeded with seed.  Returns 1
 * on a collision; orient_out / mu_out (may be NULL) receive the orientation
 * index drawn and, on a collision, the mu used; vn_out the neutral. */
int oracle_mt_ehss_offline(uint64_t seed, const icarion_domain_desc *domain, uint32_t n_orient,
                           uint32_t mu_stride, const double *sigma_eff, const double *mu,
                           double mass_kg, double vel[3], double dt, uint32_t *orient_out,
                           double *mu_out, double vn_out[3]);

/* ---- EHSS offline precompute (tools/ehss_offline_precompute.cpp:438-517) ----
 * Per (cluster c, orientation k): Shoemake orientation, b_max = max(r_xy + r) +
 * r_gas, sigma_eff = hits/n_sigma * pi b_max^2, then n_mu first-contact samples
 * mu = -n_z conditional on a hit (at most max_tries rays).  Atoms are float
 * (x, y, z, r) in metres; cluster c uses atoms [offsets[c], offsets[c] + counts[c]).
 * Random numbers: the Philox stream (seed ^ ORACLE_PRECOMPUTE_KEY, cluster,
 * orientation) the device kernel draws.  mu rows of clusters that ran out of
 * tries are filled with NaN. */
#define ORACLE_PRECOMPUTE_KEY 0x6a09e667f3bcc909ull
void oracle_ehss_precompute(const float *atoms_xyzr, const uint32_t *offsets,
                            const uint32_t *counts, uint32_t n_clusters, double r_gas,
                            uint32_t n_orient, uint32_t n_sigma, uint32_t n_mu,
                            uint32_t max_tries, uint64_t seed, double *sigma_out,
                            double *mu_out);

/* Cluster shape features of the surrogate input:
 * f1 = ln(n_atoms)/ln(500), f2 = R_g / sqrt(sigma_oapa/pi),
 * f3 = (mean atom radius + r_gas) / sqrt(sigma_oapa/pi). */
void oracle_ehss_features(const float *atoms_xyzr, const uint32_t *offsets,
                          const uint32_t *counts, uint32_t n_clusters, double r_gas,
                          const double *sigma_oapa, float *features_out);

/* OAPA of many clusters given as float atoms (oracle_oapa_sigma per cluster). */
void oracle_oapa_sigma_f32(const float *atoms_xyzr, const uint32_t *offsets,
                           const uint32_t *counts, uint32_t n_clusters, double r_gas,
                           double *sigma_out);

/* ---- surrogate MLP 6 -> 128 -> 128 -> 3 (ReLU), parameters in the layout
 * W1[128][6], b1[128], W2[128][128], b2[128], W3[3][128], b3[3] ---------- */
#define ORACLE_MLP_NPARAMS 17795
void oracle_mlp_forward(const float *params, const float *X, size_t n, float *Y);
/* Mean over the batch of 0.5*|y - t|^2 and its gradient (grad has
 * ORACLE_MLP_NPARAMS entries, overwritten). */
double oracle_mlp_loss_grad(const float *params, const float *X, const float *T, size_t n,
                            double *grad);

#ifdef __cplusplus
}
#endif

#endif
