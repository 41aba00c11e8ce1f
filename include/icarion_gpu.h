/*
 * icarion_gpu.h - C ABI of the MI355X (gfx950) batch ion-step library.
 *
 * One handle owns device-resident ion state for one ion cloud in one domain
 * and advances it by whole engine steps:
 *   prepare (birth, domain lookup, kill) -> push (RK4 or Boris) ->
 *   collision (HSS / EHSS / OU thermal kick) -> finalize (boundary
 *   elimination, NaN kill, t += dt).
 * The reference engine step this replaces is
 *   src/core/integrator/SimulationEngine.cpp:374-428 (process_timestep),
 *   src/core/integrator/SimulationEngineLifecycle.cpp:72-195.
 *
 * Conventions
 *  - Every entry point returns an int status (ICARION_OK = 0).  No C++
 *    exception crosses this boundary; the message of the last failure on the
 *    calling thread is read with icarion_gpu_last_error().
 *  - All pointers are plain host pointers unless a name says "device".
 *  - SI units throughout (m, s, kg, C, V, K, Pa).
 *  - A handle is not thread-safe: one host thread drives one handle.
 *  - Nothing here allocates on the device except icarion_gpu_create(),
 *    icarion_gpu_set_species(), icarion_gpu_upload_field_array() and
 *    icarion_gpu_upload_state() (which (re)size the resident buffers).
 */
#ifndef ICARION_GPU_H
#define ICARION_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ICARION_GPU_ABI_VERSION 6

/* ---- status codes ------------------------------------------------------ */
enum {
    ICARION_OK = 0,
    ICARION_ERR_INVALID = 1,      /* bad argument or inconsistent descriptor  */
    ICARION_ERR_HIP = 2,          /* a HIP runtime call failed                */
    ICARION_ERR_UNSUPPORTED = 3,  /* configuration the device path does not
                                     cover; the caller falls back to the CPU
                                     path (reference step_batch returns false,
                                     IIntegrationStrategy.h:106-132)          */
    ICARION_ERR_STATE = 4,        /* call order violated (e.g. step before
                                     upload)                                   */
    ICARION_ERR_NO_DEVICE = 5     /* no HIP device visible                     */
};

/* ---- enumerations (names follow the reference config schema) ----------- */
/* DomainConfig::instrument (src/core/config/types/InstrumentTypes.h) */
enum icarion_instrument {
    ICARION_INSTRUMENT_IMS = 0,
    ICARION_INSTRUMENT_TIMS = 1,        /* analytic IMS field (AnalyticalFieldModel.cpp:72-73) */
    ICARION_INSTRUMENT_QUADRUPOLE = 2,  /* "Quadrupole" / QuadrupoleRF */
    ICARION_INSTRUMENT_ORBITRAP = 3
};

/* simulation.integrator */
enum icarion_integrator {
    ICARION_INTEGRATOR_RK4 = 0,    /* RK4Strategy.cpp:181-361 */
    ICARION_INTEGRATOR_BORIS = 1,  /* BorisStrategy.cpp:18-98 */
    ICARION_INTEGRATOR_RK45 = 2    /* RK45Strategy::step (RK45Strategy.cpp:354-520):
                                      Dormand-Prince 5(4), per-ion adaptive dt,
                                      engine time = max of the ion times */
};

/* physics.collision_model */
enum icarion_collision_model {
    ICARION_COLLISION_NONE = 0,
    ICARION_COLLISION_HSS = 1,       /* HSSCollisionHandler.cpp:257-376            */
    ICARION_COLLISION_EHSS = 2,      /* EHSSCollisionHandler.cpp:131-455           */
    ICARION_COLLISION_LANGEVIN = 3,  /* DampingForce(Langevin) in the push
                                        (DampingForce.cpp:150-177) + optional OU
                                        thermal kicks (CollisionKernels.cpp:171-214) */
    ICARION_COLLISION_FRICTION = 4   /* DampingForce(Friction) (DampingForce.cpp:179-272) */
};

/* arithmetic of the push */
enum icarion_precision {
    ICARION_PRECISION_FP64 = 0,       /* plain method, FP64 throughout            */
    ICARION_PRECISION_FP32_KAHAN = 1  /* FP32 state with Kahan-compensated
                                         position; FP64 time / RF phase           */
};

/* time dependence of one field-array term
 * (FieldsConfig::FieldArrayTerm::ScaleKind; CompositeFieldProvider.h:117-147) */
enum icarion_field_scale {
    ICARION_FIELD_SCALE_CONSTANT = 0,   /* constant_scale                      */
    ICARION_FIELD_SCALE_DC_AXIAL = 1,   /* the domain's dc_axial_V             */
    ICARION_FIELD_SCALE_DC_QUAD = 2,    /* the domain's dc_quad_V              */
    ICARION_FIELD_SCALE_DC_RADIAL = 3,  /* the domain's dc_radial_V            */
    ICARION_FIELD_SCALE_RF = 4          /* rf_voltage_V * cos(2 pi f t + phase),
                                           f = frequency_Hz, or the domain's
                                           rf_frequency_Hz when that is 0     */
};

/* Waveform of a time-varying voltage (ValueOrWaveform; WaveformConfig.h:19-200).
 * Parameters p[] by type:
 *   CONSTANT     p0 value
 *   LINEAR       p0 start value, p1 end value, p2 start time, p3 end time; clamp
 *   QUADRATIC    p0 a, p1 b, p2 c (a + b t + c t^2), p3 start time, p4 end time
 *   EXPONENTIAL  p0 offset, p1 amplitude, p2 rate [1/s], p3 start, p4 end; clamp
 *   SINUSOIDAL   p0 offset, p1 amplitude, p2 frequency [Hz], p3 phase [rad]
 *   PWM          p0 low, p1 high, p2 frequency, p3 duty cycle, p4 phase, p5 start,
 *                p6 end; clamp
 *   PULSED       p0 low, p1 high, p2 pulse start, p3 pulse width
 *   ARBITRARY    times_s / values (n_points, times increasing), interp */
enum icarion_waveform_type {
    ICARION_WAVEFORM_CONSTANT = 0,
    ICARION_WAVEFORM_LINEAR = 1,
    ICARION_WAVEFORM_QUADRATIC = 2,
    ICARION_WAVEFORM_EXPONENTIAL = 3,
    ICARION_WAVEFORM_SINUSOIDAL = 4,
    ICARION_WAVEFORM_PWM = 5,
    ICARION_WAVEFORM_PULSED = 6,
    ICARION_WAVEFORM_ARBITRARY = 7
};

enum icarion_waveform_interp {
    ICARION_INTERP_LINEAR = 0,
    ICARION_INTERP_STEP = 1,
    ICARION_INTERP_CUBIC = 2     /* Catmull-Rom Hermite (WaveformConfig.cpp:54-85) */
};

/* The voltages of the analytic fields that may follow a waveform. */
enum icarion_voltage_slot {
    ICARION_SLOT_DC_AXIAL = 0,
    ICARION_SLOT_DC_QUAD = 1,
    ICARION_SLOT_DC_RADIAL = 2,
    ICARION_SLOT_RF_VOLTAGE = 3,
    ICARION_SLOT_RF_FREQUENCY = 4,
    ICARION_N_VOLTAGE_SLOTS = 5
};

/* how an EHSS collision picks the scattering normal */
enum icarion_ehss_mode {
    ICARION_EHSS_GEOMETRY = 0,   /* first-contact ray test over the atoms      */
    ICARION_EHSS_SURROGATE = 1   /* MLP surrogate (icarion_gpu_set_surrogate)  */
};

/* ---- descriptors -------------------------------------------------------- */

/* One simulation domain (DomainConfig: geometry + fields + env).  Values are
 * the evaluated constants the reference loader produces
 * (DomainConfigLoader.cpp:413-440 for E/N -> axial_V,
 *  EnvironmentConfig.h:150-282 for n, gas mass, radius, polarizability). */
typedef struct icarion_domain_desc {
    int32_t instrument;           /* enum icarion_instrument                  */
    int32_t use_field_array;      /* 1: E from the uploaded field array
                                     (GridFieldProvider) instead of the
                                     analytic model                           */
    int32_t magnetic_enabled;     /* 1: uniform B_T used by Boris             */
    int32_t reserved0;
    double origin_m[3];
    double length_m;
    double radius_m;              /* cylinder radius / quadrupole r0          */
    double radius_in_m;           /* Orbitrap inner electrode R1              */
    double radius_out_m;          /* Orbitrap outer electrode R2              */
    double radius_char_m;         /* Orbitrap characteristic radius R_m       */
    double rot_g2l[9];            /* row-major global->local rotation          */
    double rot_l2g[9];            /* row-major local->global rotation          */
    /* analytic field inputs (an absent value is 0) */
    double dc_axial_V;
    double dc_quad_V;
    double dc_radial_V;
    double rf_voltage_V;
    double rf_frequency_Hz;
    double field_array_scale;     /* factor on the interpolated field (1.0 for
                                     a single field_array_terms entry,
                                     PhysicsSetup.cpp:168-171)                */
    double B_T[3];
    /* environment */
    double temperature_K;
    double pressure_Pa;
    double gas_density_m3;        /* P/(kB T)                                 */
    double gas_mass_kg;
    double gas_radius_m;
    double gas_polarizability_m3;
    double gas_velocity_m_s[3];
    /* arrival detection: global z of the detector plane; NaN disables */
    double detector_z_m;
} icarion_domain_desc;

/* physics switches for one handle */
typedef struct icarion_physics_desc {
    int32_t integrator;           /* enum icarion_integrator                  */
    int32_t collision_model;      /* enum icarion_collision_model             */
    int32_t ou_thermal_kicks;     /* Langevin/Friction: add OU thermal kicks
                                     with the ion's own gamma (apply_damping =
                                     false, CollisionKernels.cpp:204-213)     */
    int32_t precision;            /* enum icarion_precision                   */
    int32_t ehss_mode;            /* enum icarion_ehss_mode                   */
    int32_t collision_substeps;   /* collision stages per engine step, each
                                     with dt / substeps on the pushed state
                                     (handle_collisions_cpu's loop,
                                     SimulationEngineCollisionRuntime.cpp:
                                     84-191): collision_subcycles_per_step, or
                                     max(it, collision_max_events_per_step)
                                     in multi-event mode (:26-33); 0 or 1 =
                                     one stage.  Not with the wave-cooperative
                                     geometry EHSS of a single-domain
                                     fixed-step handle or the EHSS surrogate */
    uint64_t rng_seed;            /* simulation.rng_seed                      */
} icarion_physics_desc;

#define ICARION_MAX_COLLISION_SUBSTEPS 1024

/* Adaptive step control of ICARION_INTEGRATOR_RK45
 * (RK45Strategy::AdaptiveConfig, RK45Strategy.h:54-65; defaults in brackets,
 * used until icarion_gpu_set_rk45 is called). */
typedef struct icarion_rk45_desc {
    double atol;                  /* absolute tolerance [m, m/s]     (1e-8)   */
    double rtol;                  /* relative tolerance              (1e-6)   */
    double safety_factor;         /*                                 (0.9)    */
    double min_step_factor;       /* dt_min = dt_ion * factor        (1e-6)   */
    double max_step_factor;       /* dt_max = dt_ion * factor        (1e3)    */
    double max_step_increase;     /* growth clamp per step           (5.0)    */
    double max_step_decrease;     /* shrink clamp per step           (0.1)    */
    double absolute_min_step_s;   /* dt floor, 0 = none              (0.0)    */
    int32_t accept_at_dt_min;     /* accept at dt_min even if err > 1 (1)     */
    int32_t reserved0;
} icarion_rk45_desc;

/* Per-species constants, resolved on the host for the domain's gas
 * (species-DB priorities: HSSCollisionHandler.cpp:263-310,
 *  EHSSCollisionHandler.cpp:462-547). */
typedef struct icarion_species_desc {
    double sigma_hss_m2;          /* HSS cross section for the domain gas     */
    double sigma_ehss_m2;         /* EHSS sigma_eff for the domain gas        */
    double ccs_m2;                /* ion CCS                                   */
    double K0_cm2_Vs;             /* reduced mobility (Friction damping)       */
    uint32_t atom_offset;         /* EHSS geometry: first atom in the table    */
    uint32_t atom_count;          /* EHSS geometry: number of atoms            */
    float surrogate_features[3];  /* cluster shape features of the EHSS
                                     surrogate (icarion_gpu_synth_clusters)    */
    uint32_t reserved0;
} icarion_species_desc;

/* One component of a gas mixture (EnvironmentConfig gas_mixture entry). */
typedef struct icarion_gas_component {
    double density_m3;            /* n_i                                       */
    double mass_kg;               /* neutral mass                              */
    int32_t participates;         /* participates_in_collisions                */
    int32_t reserved0;
} icarion_gas_component;

#define ICARION_MAX_GAS_COMPONENTS 4

/* Where a domain's ValueOrWaveform takes its waveform from (WaveformConfig.h:
 * 242-277).  The analytic fields evaluate both alike.  The time scales of a
 * composite field array (two or more terms) resolve only constants and library
 * references, and take an inline waveform as 0 (CompositeFieldProvider.h:150-171). */
enum icarion_waveform_source {
    ICARION_WAVEFORM_LIBRARY = 0,   /* waveform_ref into the domain's waveform_library */
    ICARION_WAVEFORM_INLINE = 1     /* an inline waveform                              */
};

typedef struct icarion_waveform_desc {
    int32_t type;                 /* enum icarion_waveform_type                */
    int32_t interp;               /* ARBITRARY: enum icarion_waveform_interp   */
    int32_t clamp;                /* LINEAR / EXPONENTIAL / PWM: hold the end
                                     value after the end time                  */
    int32_t source;               /* enum icarion_waveform_source              */
    double p[8];                  /* parameters, see icarion_waveform_type     */
    const double *times_s;        /* ARBITRARY                                 */
    const double *values;         /* ARBITRARY                                 */
    uint32_t n_points;            /* ARBITRARY                                 */
    uint32_t reserved1;
} icarion_waveform_desc;

/* Surrogate training options (Adam on mean 0.5*|y - t|^2). */
typedef struct icarion_train_desc {
    uint32_t epochs;
    uint32_t batch;               /* samples per step                          */
    float lr;
    float beta1, beta2, eps;
    uint64_t seed;                /* initialisation and shuffling              */
} icarion_train_desc;

#define ICARION_SURROGATE_NPARAMS 17795

/* Host SoA view of an ion ensemble (IonEnsemble hot/cold blocks,
 * src/core/types/IonEnsemble.h:267-318).  Inputs are read by upload, outputs
 * written by download; an output pointer may be NULL to skip it. */
typedef struct icarion_ion_soa {
    size_t n;
    double *x, *y, *z;            /* position [m]                (in/out)     */
    double *vx, *vy, *vz;         /* velocity [m/s]              (in/out)     */
    double *mass_kg;              /* (in)                                      */
    double *charge_C;             /* (in)                                      */
    uint32_t *species;            /* index into the species table (in)         */
    uint8_t *active;              /* (in/out)                                  */
    uint8_t *born;                /* (in/out)                                  */
    double *birth_time_s;         /* (in; NULL = born at upload time)          */
    double *death_time_s;         /* (in/out; -1 while alive)                  */
    double *arrival_time_s;       /* (out; first detector-plane crossing,
                                     NaN if none)                              */
    uint32_t *n_collisions;       /* (out; collision events so far)            */
    double *time_s;               /* (in/out) ion time t_i (IonEnsemble output
                                     block); RK45 advances it by each ion's
                                     accepted dt.  NULL or NaN on upload = the
                                     t0 of the first step                      */
    double *dt_s;                 /* (in/out) RK45 per-ion dt (the engine's
                                     dt_per_ion_); NULL or <= 0 on upload = the
                                     dt of the first step                      */
} icarion_ion_soa;

/* Counters of one icarion_gpu_step() call. */
typedef struct icarion_step_stats {
    uint64_t ion_steps;           /* sum over steps of ions pushed             */
    uint64_t collisions;          /* collision events                          */
    uint64_t alive;               /* active and born after the call            */
    uint64_t arrived;             /* ions with an arrival time after the call  */
    uint64_t rejected;            /* RK45: rejected trial steps                */
    double t_end_s;               /* engine time after the call                */
    double device_ms;             /* device time of the stepping kernels       */
} icarion_step_stats;

/* handle creation options */
typedef struct icarion_gpu_desc {
    int32_t device;               /* HIP device ordinal                        */
    uint32_t block_size;          /* threads per workgroup; 0 = tuned default  */
    uint32_t steps_per_launch;    /* steps fused per kernel launch; 0 = default */
    uint32_t reserved0;
} icarion_gpu_desc;

typedef struct icarion_gpu icarion_gpu;

/* ---- entry points -------------------------------------------------------- */

/* Library ABI version (ICARION_GPU_ABI_VERSION of the build). */
int icarion_gpu_abi_version(void);

/* Number of HIP devices visible; 0 if none. */
int icarion_gpu_device_count(void);

/* Message of the last failed call on this thread; returns its full length. */
size_t icarion_gpu_last_error(char *buf, size_t len);

/* Create / destroy a handle.  Replaces GPUIntegrationHelper / GPUCollisionHelper
 * construction (src/core/integrator/strategies/GPUIntegrationStrategy.cpp:100-140). */
int icarion_gpu_create(const icarion_gpu_desc *desc, icarion_gpu **out);
int icarion_gpu_destroy(icarion_gpu *h);

/* Domain + physics of the handle.  Mirrors what IntegrationStrategyFactory::create
 * and CollisionHandlerFactory::create receive (IntegrationStrategyFactory.h:27-47,
 * CollisionHandlerFactory.cpp:10-18). */
int icarion_gpu_set_domain(icarion_gpu *h, const icarion_domain_desc *domain,
                           const icarion_physics_desc *physics);

/* Species table and EHSS atom table (x, y, z, r per atom, metres, COM-centred,
 * geometryUtils.h:35-60).  atoms_dtype 0 = float, 1 = double.  The device
 * keeps the table in float and the EHSS geometry arithmetic runs in double on
 * those values; a double table (atoms_dtype 1 with n_atoms > 0) returns
 * ICARION_ERR_UNSUPPORTED, so that the caller rounds it or keeps the CPU path.
 * atoms_xyzr may be NULL when n_atoms is 0. */
int icarion_gpu_set_species(icarion_gpu *h, const icarion_species_desc *species,
                            uint32_t n_species, const void *atoms_xyzr,
                            uint32_t n_atoms, int32_t atoms_dtype);

/* HSS over a gas mixture: HSSCollisionHandler's mixture path
 * (HSSCollisionHandler.cpp:105-249) - per participating component with
 * n_i > 0 a Maxwellian neutral at the domain temperature and flow, rate
 * k_i = n_i sigma_i |v - v_i| (skipped below the minimum relative speed), one
 * collision test on k_total, the partner drawn with weight k_i, an HSS
 * collision with that neutral.  sigma_m2[s * n_components + c] is species s's
 * cross section with component c, resolved by the caller with the handler's
 * priority (species-DB CCS_HSS for that gas > derived from the reference
 * gas > the component's cross_section_m2 > the ion CCS); it must be > 0 for
 * every component that participates with n_i > 0 (the reference throws
 * otherwise).  n_species must equal that of icarion_gpu_set_species.
 * n_components = 0 returns to the single-gas path.
 * With EHSS (geometry method) the same table holds the EHSS effective cross
 * sections per gas (compute_effective_ccs with the component's radius), and
 * the handler's mixture path runs (EHSSCollisionHandler.cpp:311-424): every
 * component visited in order, a neutral drawn for each participating one
 * (also at n_i = 0), k_i = n_i sigma_i |v - v_i| where all three are > 0,
 * one test on k_total, the partner drawn with weight k_i (the last component
 * when the running sum falls short), then the first-contact geometry method
 * with that neutral, mass and contact radius (icarion_gpu_set_gas_radii).
 * Offline sample sets are not used with a mixture, as in the reference; EHSS
 * orientation samples and the surrogate return ICARION_ERR_UNSUPPORTED with it. */
int icarion_gpu_set_gas_mixture(icarion_gpu *h, const icarion_gas_component *components,
                                uint32_t n_components, const double *sigma_m2, uint32_t n_species);

/* Contact radius of each gas component of the mixture (GasMixtureComponent
 * radius_m; 0 = the domain's gas_radius_m), for EHSS over a mixture. */
int icarion_gpu_set_gas_radii(icarion_gpu *h, const double *radius_m, uint32_t n_components);

/* EHSS orientation samples per species (EHSSOrientationSamples,
 * EHSSSamples.h:14-17: orientations_quat and the areas_by_gas_m2 entry of the
 * domain's gas), as EHSSCollisionHandler::handle_collision uses them
 * (EHSSCollisionHandler.cpp:187-217, 286-293).  A species with n_orient[s] > 0
 * and no offline sample set draws one of its orientations uniformly, takes its
 * area as sigma_eff in the collision test and, on a collision, runs the
 * first-contact test with the molecule in that orientation
 * (CollisionKernels::ehss_collision_with_orientation with that sigma_eff and
 * the forced-hit budget of 16 x 256 impact parameters).  quat_wxyz: 4 values
 * per orientation (w, x, y, z; normalised on use as quaternion_to_rotation
 * does), area_m2: one per orientation, both species-major.  n_species must
 * equal that of icarion_gpu_set_species; n_species = 0 clears.  EHSS geometry
 * method, FP64. */
int icarion_gpu_set_ehss_orientations(icarion_gpu *h, const uint32_t *n_orient, const double *quat_wxyz,
                                      const double *area_m2, uint32_t n_species);

/* Time-varying voltages of a domain: slots[k] (ICARION_N_VOLTAGE_SLOTS
 * entries, NULL = the domain's constant) replaces the value of slot k,
 * evaluated at each stage time.
 *   - Analytic IMS / TIMS and quadrupole fields, as
 *     AnalyticalFieldModel::compute_ims_field / compute_quadrupole_rf_field do
 *     (AnalyticalFieldModel.cpp:130-248: E_z = axial(t)/L, U_eff = quad(t) +
 *     V(t) cos(2 pi f(t) t) where |V(t)| > 1e-12).
 *   - The term scales of a composite field array (two or more terms), as
 *     CompositeFieldProvider::compute_scale_factor does
 *     (CompositeFieldProvider.h:117-171: DC kinds V(t); RF V(t) cos(2 pi f t +
 *     phi) with f the term's own frequency, else f(t); an inline waveform
 *     counts as 0 there).  One field array is not scaled (GridFieldProvider).
 *   - The Orbitrap: compute_orbitrap_field reads the DC radial slot at t = 0
 *     only (AnalyticalFieldModel.cpp:188-202), so its waveform fixes the field
 *     constant at V(0) and the other slots are not read; any integrator.
 * The fixed-step integrators (RK4, Boris) take them from FP64 tables the host
 * evaluates per step; RK45 evaluates them in the kernel at each ion's own
 * stage times.  slots = NULL clears.  RK45 on a field array keeps constant
 * voltages (ICARION_ERR_UNSUPPORTED when combined).  Batched clouds take
 * theirs in icarion_cloud_desc.waveforms. */
int icarion_gpu_set_waveforms(icarion_gpu *h, const icarion_waveform_desc *const *slots);

/* The gas pressure of the domain as a waveform (EnvironmentConfig::
 * pressure_Pa_waveform, EnvironmentConfig.h:58; any waveform there counts as
 * time-varying, a CONSTANT one included).  Replaces
 * SimulationEngine::update_dynamic_environments (SimulationEngineLifecycle.cpp:
 * 34-47), which sets P(t) at the engine time t of each step start and derives
 * n = P(t)/(kB T) and a mixture's n_i = x_i n (EnvironmentConfig.h:132-175):
 * every rate of that step (each gas component's collision rate, the Langevin
 * and Friction damping coefficients and with them the OU kicks) scales by
 * n / gas_density_m3 of the domain descriptor, which must be > 0 (it holds
 * the constants the rest derive from: the reference's P(0)).  A step whose
 * P(t) <= 0 returns ICARION_ERR_INVALID, as the reference throws (the steps
 * before it have run).  RK4, Boris and RK45 (evaluated in the kernel at each
 * engine step's time), FP64 and FP32, the EHSS surrogate included.
 * pressure = NULL clears. */
int icarion_gpu_set_pressure_waveform(icarion_gpu *h, const icarion_waveform_desc *pressure);

/* Gas-flow profiles of EnvironmentConfig (flow_model, EnvironmentConfig.h:22-26,
 * 60-62).  CONSTANT and AxialUniform are a constant velocity, which the domain
 * descriptor's gas_velocity_m_s carries (AxialUniform as R_l2g (0, 0, u)). */
enum icarion_flow_model {
    ICARION_FLOW_CONSTANT = 0,          /* the descriptor's gas_velocity_m_s    */
    ICARION_FLOW_AXIAL_PARABOLIC = 1    /* Poiseuille profile, below            */
};

/* The domain's gas flow as EnvironmentConfig::gas_velocity_at gives it
 * (EnvironmentConfig.h:96-129), evaluated at each ion's own position: for
 * ICARION_FLOW_AXIAL_PARABOLIC, u = R_l2g (0, 0, u_max max(0, 1 - r^2 / R^2))
 * with r the radial distance in the domain's local frame (R_g2l (p - origin))
 * and R the descriptor's radius_m (radius_m <= 0: u_max everywhere).  The
 * engine reads it in the damping force at every force evaluation
 * (DampingForce.cpp:83-88, stage positions) and for the neutral velocity of
 * every collision substep at the ion's pushed position
 * (SimulationEngineCollisionRuntime.cpp:122-130).  Fixed-step integrators,
 * FP64 and FP32, HSS (single gas and mixture), EHSS with offline sets,
 * orientation samples or a mixture, Langevin / Friction, and RK45 (its
 * damping force at each stage position); the wave-cooperative geometry EHSS
 * kernel and the EHSS surrogate return ICARION_ERR_UNSUPPORTED with it
 * (several domains take theirs with icarion_gpu_set_domain_gas_flow).
 * ICARION_FLOW_CONSTANT clears. */
int icarion_gpu_set_gas_flow(icarion_gpu *h, int32_t flow_model, double axial_max_velocity_m_s);

/* Offline EHSS sample sets (EHSSOfflineSampleSet, EHSSOfflineSampleSet.h:15-30;
 * what tools/ehss_offline_precompute.cpp and icarion_gpu_ehss_precompute
 * produce).  Species s gets n_orient[s] orientations (0: it keeps the online
 * geometry method), each with a sigma_eff and mu_stride[s] first-contact mu;
 * sigma_eff and mu hold the sets of all species concatenated in species order
 * (mu orientation-major).  An EHSS collision of such a species draws an
 * orientation first, tests with its sigma_eff and scatters with one of its mu
 * at a uniform azimuth (EHSSCollisionHandler.cpp:151-172, 222-299).  Used by
 * the FP64 EHSS path (ICARION_EHSS_GEOMETRY, fixed-step and RK45); n_species
 * must equal that of icarion_gpu_set_species; n_species = 0 clears. */
int icarion_gpu_set_ehss_offline(icarion_gpu *h, const uint32_t *n_orient, const uint32_t *mu_stride,
                                 const double *sigma_eff, const double *mu, uint32_t n_species);

/* Trilinear E field on a uniform node-centred grid, x fastest
 * (index i + nx*(j + ny*k), fieldArrayLoader.cpp:9-11).  E holds 3*nx*ny*nz
 * values interleaved (Ex, Ey, Ez) per node; dtype 0 = float, 1 = double.
 * Axes: x_i = x0 + i*dx etc.  Replaces load_field_array + GridFieldProvider
 * (fieldArrayLoader.cpp:13-162, GridFieldProvider.h:34-98). */
int icarion_gpu_upload_field_array(icarion_gpu *h, uint32_t nx, uint32_t ny, uint32_t nz,
                                   double x0, double dx, double y0, double dy,
                                   double z0, double dz, const void *E, int32_t dtype);

/* One entry of a domain's field_array_terms: a field array on the shared
 * grid of icarion_gpu_upload_field_terms and its scaling. */
typedef struct icarion_field_term_desc {
    const void *E;                /* 3*nx*ny*nz values, (Ex, Ey, Ez) per node,
                                     x fastest                                */
    int32_t dtype;                /* 0 = float, 1 = double                     */
    int32_t scale_kind;           /* enum icarion_field_scale                  */
    double constant_scale;        /* ICARION_FIELD_SCALE_CONSTANT              */
    double frequency_Hz;          /* ICARION_FIELD_SCALE_RF (0 = domain's)     */
    double phase_rad;             /* ICARION_FIELD_SCALE_RF                    */
} icarion_field_term_desc;

#define ICARION_MAX_FIELD_TERMS 4

/* The field-array source of a domain with use_field_array = 1, on a
 * rectilinear node-centred grid whose axes may be non-uniform (xs[nx], ys[ny],
 * zs[nz], non-decreasing).  Interpolation is trilinear with clamping outside
 * the grid, the node of each axis found as by the reference's axis search
 * (interpolate_field, fieldArrayLoader.cpp:103-162).  One term is used
 * unscaled (GridFieldProvider, PhysicsSetup.cpp:165-171); 2 to
 * ICARION_MAX_FIELD_TERMS terms are superposed as sum_i scale_i(t) E_i(r)
 * (CompositeFieldProvider::get_E, CompositeFieldProvider.h:80-97), with the
 * voltages taken from the domain set by icarion_gpu_set_domain at step time.
 * The terms share the grid (unit-voltage solutions of one mesh); the domain's
 * field_array_scale multiplies the result.  Replaces load_field_array +
 * GridFieldProvider / CompositeFieldProvider (PhysicsSetup.cpp:107-176). */
int icarion_gpu_upload_field_terms(icarion_gpu *h, uint32_t nx, uint32_t ny, uint32_t nz,
                                   const double *xs, const double *ys, const double *zs,
                                   const icarion_field_term_desc *terms, uint32_t n_terms);

/* Copy host SoA state to the device (resident until the next upload). */
int icarion_gpu_upload_state(icarion_gpu *h, const icarion_ion_soa *ions);

/* Advance the resident ensemble by n_steps engine steps of dt starting at
 * engine time t0.  step_offset numbers the first step for the per-ion RNG
 * streams (Philox4x32-10 keyed by (rng_seed, ion index), counter = step).
 * stats may be NULL.  The device path of IIntegrationStrategy::step_batch +
 * ICollisionHandler::handle_batch + finalize_ions_after_integration, fused. */
int icarion_gpu_step(icarion_gpu *h, double t0, double dt, uint64_t n_steps,
                     uint64_t step_offset, icarion_step_stats *stats);

/* icarion_gpu_step that allocates nothing, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The step table
 * and launch counters live in a fixed scratch array of the library (one per
 * device, 1 MiB + 8 KiB, held until the process ends) instead of the handle's
 * buffers, which icarion_gpu_step resizes to each call.  Limits: at most
 * ICARION_NOALLOC_MAX_STEPS steps per call (ICARION_ERR_INVALID above);
 * single-domain handles with RK4 or Boris (ICARION_ERR_UNSUPPORTED for
 * several domains, RK45 and waveform-scaled field-array terms).  The ion
 * state, species and field buffers are the handle's, set up beforehand.
 * Results equal icarion_gpu_step's bit for bit.  Every call on one device
 * shares that device's scratch: calls that overlap on one device (two host
 * threads, two handles) race and must be serialised by the caller. */
#define ICARION_NOALLOC_MAX_STEPS 16384
int icarion_gpu_step_noalloc(icarion_gpu *h, double t0, double dt, uint64_t n_steps,
                             uint64_t step_offset, icarion_step_stats *stats);

/* Copy the resident state back into host SoA arrays. */
int icarion_gpu_download_state(icarion_gpu *h, icarion_ion_soa *ions);

/* Arrival-time statistics of the resident ions: the finite in-kernel
 * arrival times (first crossing of the detector plane, linearly interpolated
 * in the step; NaN = none). */
typedef struct icarion_arrival_stats {
    uint64_t count;
    double min_s, max_s;
    double mean_s, std_s;         /* population std, as numpy's values.std()  */
} icarion_arrival_stats;

#define ICARION_MAX_HIST_BINS 8192

/* The arrival statistics (analysis/ims_mobility.py and arrival_time_distribution.py
 * take the mean and std of the arrival times), reduced on the device per
 * workgroup and summed on the host in a fixed order. */
int icarion_gpu_arrival_stats(icarion_gpu *h, icarion_arrival_stats *stats);

/* Counts of the arrival times in n_bins equal bins over [lo_s, hi_s] with
 * numpy's bin rule (np.histogram over the edges linspace(lo, hi, n_bins + 1):
 * each bin [e_i, e_i+1), the last one closed), as the reference's analysis
 * bins them with edges linspace(min, max, bins + 1) (analysis/physics.py:96-104,
 * 119-145).  1 <= n_bins <= ICARION_MAX_HIST_BINS, lo_s < hi_s. */
int icarion_gpu_arrival_histogram(icarion_gpu *h, double lo_s, double hi_s, uint32_t n_bins, uint64_t *counts);

#define ICARION_MAX_DOMAINS 8

/* Several domains on one handle: a multi-domain instrument as DomainManager
 * holds it (DomainManager.cpp:16-229) with the engine's per-step domain logic
 * (SimulationEngineLifecycle.cpp:72-195):
 *   - at each step start an ion takes the first domain that contains it, by
 *     find_domain_index's z-sorted span search and then its linear scan, and
 *     dies if none does;
 *   - it moves in that domain's field and collides with that domain's gas;
 *   - after the push, an ion in no domain lives on only when it left its
 *     domain through the end face into the next one
 *     (forward_axial_bridge_domain); its domain index follows
 *     shared_boundary_handoff_domain, with end aperture end_aperture_m[k]
 *     (0 or NULL: the radius).
 * Replaces any domain set before; n_domains = 1 is icarion_gpu_set_domain.
 * Covered: analytic fields (IMS / TIMS, quadrupole, Orbitrap) with constant
 * voltages or per-domain waveforms (icarion_gpu_set_domain_waveforms), uniform
 * field arrays per domain (icarion_gpu_upload_domain_field_array), RK4 or Boris, FP64 or FP32 + Kahan (domain lookup, bridge and
 * handoff in FP64 on the compensated position), no / HSS / Langevin /
 * Friction collisions (HSS also with per-domain gas mixtures,
 * icarion_gpu_set_domain_gas_mixture) and geometry EHSS (FP64, online), the
 * handle's species table or a domain's own (icarion_gpu_set_domain_species);
 * icarion_gpu_step, icarion_gpu_step_batch and
 * icarion_gpu_step_snapshots (with each ion's domain index).  Other
 * combinations return ICARION_ERR_UNSUPPORTED when stepped. */
int icarion_gpu_set_domains(icarion_gpu *h, const icarion_domain_desc *domains,
                            const double *end_aperture_m, uint32_t n_domains,
                            const icarion_physics_desc *physics);

/* The field array of one domain of a multi-domain set whose descriptor has
 * use_field_array = 1: that domain's GridFieldProvider, a uniform grid with
 * one unscaled term (times the domain's field_array_scale), with the
 * arguments and rules of icarion_gpu_upload_field_array.  An ion in that
 * domain moves in its array (clamped trilinear interpolation), elsewhere in
 * its domain's analytic field.  icarion_gpu_set_domains clears every domain's
 * array.  No reference counterpart as a call: it carries the field_array_terms
 * of one entry of FullConfig::domains (PhysicsSetup.cpp:96-249). */
int icarion_gpu_upload_domain_field_array(icarion_gpu *h, uint32_t domain, uint32_t nx, uint32_t ny, uint32_t nz,
                                          double x0, double dx, double y0, double dy, double z0, double dz,
                                          const void *E_xyz, int32_t dtype);

/* Voltage waveforms of one domain of a multi-domain set (each domain's
 * ValueOrWaveform fields and waveform_library, DomainConfig.h), with the
 * slots and rules of icarion_gpu_set_waveforms for that domain's analytic
 * field: IMS / TIMS and quadrupole voltages evaluated at every stage time, the
 * Orbitrap's DC radial voltage at t = 0.  slots = NULL clears the domain's
 * waveforms; icarion_gpu_set_domains clears every domain's.  Handle-level
 * waveforms (icarion_gpu_set_waveforms) are not combined with several domains:
 * stepping returns ICARION_ERR_UNSUPPORTED.  No reference counterpart as a
 * call: it carries the per-domain part of FullConfig::domains. */
int icarion_gpu_set_domain_waveforms(icarion_gpu *h, uint32_t domain, const icarion_waveform_desc *const *slots);

/* The pressure waveform of one domain of a multi-domain set (each domain's
 * environment, updated at every step start by update_dynamic_environments,
 * SimulationEngineLifecycle.cpp:34-47), with the rules of
 * icarion_gpu_set_pressure_waveform for that domain's gas; NULL clears.
 * icarion_gpu_set_domains clears every domain's; a handle-level pressure
 * waveform is not combined with several domains (ICARION_ERR_UNSUPPORTED). */
int icarion_gpu_set_domain_pressure_waveform(icarion_gpu *h, uint32_t domain, const icarion_waveform_desc *pressure);

/* The gas flow of one domain of a multi-domain set (its EnvironmentConfig's
 * flow_model, read with that domain's geometry and rotations as the engine
 * passes them, SimulationEngineCollisionRuntime.cpp:93-101), with the rules of
 * icarion_gpu_set_gas_flow: an ion takes its step domain's flow.
 * icarion_gpu_set_domains clears every domain's; a handle-level flow is not
 * combined with several domains (ICARION_ERR_UNSUPPORTED). */
int icarion_gpu_set_domain_gas_flow(icarion_gpu *h, uint32_t domain, int32_t flow_model,
                                    double axial_max_velocity_m_s);

/* The HSS gas mixture of one domain of a multi-domain set (that domain's
 * EnvironmentConfig gas_mixture), with the arguments and rules of
 * icarion_gpu_set_gas_mixture; an ion collides with its step domain's mixture,
 * or with its single gas where the domain has none.  HSS and geometry EHSS
 * (FP64; every component's contact radius is the domain's gas radius).  n_components = 0 clears
 * the domain's; icarion_gpu_set_domains clears every domain's.  A handle-level
 * mixture (icarion_gpu_set_gas_mixture) is not combined with several domains.
 * No reference counterpart as a call: it carries the per-domain part of
 * FullConfig::domains. */
int icarion_gpu_set_domain_gas_mixture(icarion_gpu *h, uint32_t domain, const icarion_gas_component *components,
                                       uint32_t n_components, const double *sigma_m2, uint32_t n_species);

/* The species table of one domain of a multi-domain set: the species of
 * icarion_gpu_set_species in the same order (n_species equal), with the cross
 * sections, CCS and K0 resolved for that domain's gas.  The reference resolves
 * them per ion species and the environment of the ion's domain
 * (HSSCollisionHandler.cpp:263-310, EHSSCollisionHandler.cpp:462-547, called
 * with the domain's EnvironmentConfig from SimulationEngineCollisionRuntime.cpp:
 * 36-46; DampingForce reads K0 and the domain's gas, DampingForce.cpp:105-277).
 * An ion takes the row of its step domain.  Each row keeps the handle's atom
 * range (one geometry per species).  species = NULL or n_species = 0 clears the
 * domain's table (the domain then uses the handle's); icarion_gpu_set_domains
 * clears every domain's.  No reference counterpart as a call: it carries the
 * per-domain part of the species-database lookup. */
int icarion_gpu_set_domain_species(icarion_gpu *h, uint32_t domain, const icarion_species_desc *species,
                                   uint32_t n_species);

/* Every resident ion's domain index after the last step (the ensemble's
 * domain_index; 0 before the ion's first step, IonState's default).  n = the
 * resident count. */
int icarion_gpu_download_domains(icarion_gpu *h, int32_t *domain_index);

/* One trajectory snapshot, in the layout OutputManager::log_step buffers and
 * HDF5Writer writes (OutputManager.cpp:197-227, hdf5Writer.cpp:158-299):
 * positions and velocities interleaved per ion [n][3] (FP64), the ion's
 * species index, domain index and flags (bit 0 active, bit 1 born).  The
 * arrays are valid during the call only.  Return 0 to continue, nonzero to
 * stop the run. */
typedef int (*icarion_snapshot_fn)(void *user, uint64_t index, uint64_t step, double t_s, size_t n,
                                   const double *positions, const double *velocities,
                                   const uint32_t *species, const int32_t *domain_index,
                                   const uint8_t *flags);

/* icarion_gpu_step with the trajectory sampling of SimulationEngine::run /
 * OutputManager (SimulationEngine.cpp:268-283, OutputManager.cpp:246-253):
 * after the first step when write_first is set, and after every step whose end
 * time t reaches *next_write_s, which then becomes t + write_interval_s.
 * Launches are cut at the snapshot steps.  Each snapshot is packed on the
 * device, and it reaches pinned host memory while the steps that follow it run.
 * After the first snapshot of a call, only the ions that may have changed
 * since the previous snapshot cross PCIe (active ones, and ones that died or
 * were born since), when they are few; fn always sees every ion.  fn is
 * called on the calling thread, in snapshot order, one snapshot behind the
 * steps queued, and it must not call into the same handle.
 * *next_write_s is updated so that a following call continues the schedule.
 * Fixed-step integrators (RK4, Boris); RK45 returns ICARION_ERR_UNSUPPORTED.
 * With no ions, fn is called with n = 0 and NULL arrays.  A nonzero return of
 * fn ends the run after the steps already queued (ICARION_ERR_STATE). */
int icarion_gpu_step_snapshots(icarion_gpu *h, double t0, double dt, uint64_t n_steps,
                               uint64_t step_offset, double write_interval_s, int write_first,
                               double *next_write_s, icarion_snapshot_fn fn, void *user,
                               icarion_step_stats *stats);

/* Step control of ICARION_INTEGRATOR_RK45 (RK45Strategy(const AdaptiveConfig&),
 * RK45Strategy.cpp:179-195: positive tolerances, safety factor in (0, 1)). */
int icarion_gpu_set_rk45(icarion_gpu *h, const icarion_rk45_desc *desc);

/* ---- EHSS geometry and the scattering surrogate ---------------------------- */

/* Build n_clusters random rigid atom clusters on the device (cluster c has
 * counts[c] atoms, radii uniform in [r_min_m, r_max_m], sequential attachment
 * at 0.95 (r_i + r_j), no closer than 0.8 (r_i + r_j), centred) and install
 * them as the species table (species c = cluster c) with sigma_ehss = the OAPA
 * area of EHSSCollisionHandler::compute_effective_ccs (EHSSCollisionHandler.cpp:474-498,
 * gas radius of the current domain) and the surrogate features.  Optional
 * host outputs: atoms (4 floats per atom, cluster-major), sigma, features (3
 * per cluster).  Needs icarion_gpu_set_domain first. */
int icarion_gpu_synth_clusters(icarion_gpu *h, uint32_t n_clusters, const uint32_t *counts,
                               double r_min_m, double r_max_m, uint64_t seed,
                               float *atoms_out, double *sigma_out, float *features_out);

/* Offline EHSS precompute (tools/ehss_offline_precompute.cpp:438-517) for
 * species [cluster_begin, cluster_begin + n_clusters) of the current atom
 * table (float atoms): per orientation sigma_eff (n_sigma rays) and n_mu
 * first-contact mu = -n.axis samples conditional on a hit (at most max_tries
 * rays; missing samples are NaN).  Outputs sigma_out[n_clusters*n_orient],
 * mu_out[n_clusters*n_orient*n_mu]. */
int icarion_gpu_ehss_precompute(icarion_gpu *h, uint32_t cluster_begin, uint32_t n_clusters,
                                uint32_t n_orient, uint32_t n_sigma, uint32_t n_mu,
                                uint32_t max_tries, uint64_t seed, double *sigma_out,
                                double *mu_out);

/* Surrogate parameters (fp32, ICARION_SURROGATE_NPARAMS, layout W1[128][6],
 * b1, W2[128][128], b2, W3[3][128], b3).  set also packs the bf16 inference
 * weights used by ICARION_EHSS_SURROGATE. */
int icarion_gpu_surrogate_set_params(icarion_gpu *h, const float *params, uint32_t n_params);
int icarion_gpu_surrogate_get_params(icarion_gpu *h, float *params, uint32_t n_params);

/* Train the surrogate on the device: X[n][6] inputs, T[n][3] targets.
 * Starts from the current parameters (or a seeded initialisation when none
 * were set); loss_per_epoch (may be NULL) receives the mean training loss of
 * each epoch.  Packs the inference weights at the end. */
int icarion_gpu_surrogate_train(icarion_gpu *h, const float *X, const float *T, size_t n,
                                const icarion_train_desc *desc, float *loss_per_epoch);

/* Loss and gradient of the current parameters on (X, T) (one fp32
 * forward + backward pass on the device; for gradient checks). */
int icarion_gpu_surrogate_grad(icarion_gpu *h, const float *X, const float *T, size_t n,
                               float *grad_out, float *loss_out);

/* Batch inference with the packed bf16 weights on the matrix cores: Y[n][3]. */
int icarion_gpu_surrogate_infer(icarion_gpu *h, const float *X, size_t n, float *Y);

/* Block until all work queued on the handle's stream (and the device) is done. */
int icarion_gpu_synchronize(icarion_gpu *h);

/* ---- batched clouds -------------------------------------------------------
 * Many independent clouds (each its own domain, physics, species and ions)
 * advanced in one persistent launch: a fixed pool of workgroups takes tiles
 * (up to block_size ions of one cloud, all of that cloud's steps) from a
 * queue ordered longest first.  Per cloud the result is that of
 * icarion_gpu_step on a handle of its own (the same per-ion code and random
 * streams; the compiler may fuse multiply-adds differently in the batch
 * kernel, so FP32 clouds can differ from a single-cloud run by FP32
 * rounding, which collisions amplify into statistical agreement).  Covered: RK4 / Boris; IMS, TIMS, quadrupole and Orbitrap
 * analytic fields; no collisions, HSS, Langevin / Friction (+ OU kicks);
 * FP64 or FP32 + Kahan per cloud; HSS gas mixtures; geometry EHSS in FP64
 * clouds (the cloud's own atom table); EHSS with the surrogate (FP32 + Kahan,
 * the handle's surrogate parameters, species sigma_ehss and features; atoms
 * are not needed); a uniform-grid field array per cloud (RK4 / Boris; no
 * collisions, HSS, Langevin / Friction).  Not covered
 * (ICARION_ERR_UNSUPPORTED): RK45, geometry EHSS in FP32 clouds, field arrays
 * with a gas mixture or EHSS, composite / non-uniform field arrays.  The batch
 * lives in the handle beside (and independent of) its single-cloud state. */
typedef struct icarion_cloud_desc {
    icarion_domain_desc domain;
    icarion_physics_desc physics;
    const icarion_species_desc *species;
    uint32_t n_species;
    uint32_t reserved0;
    icarion_ion_soa *ions;        /* host SoA: read by upload, written by
                                     download                                  */
    double t0_s;                  /* engine time of the first step             */
    double dt_s;
    uint64_t n_steps;             /* < 2^32                                    */
    uint64_t step_offset;         /* first RNG step number                     */
    const icarion_waveform_desc *const *waveforms;
                                  /* NULL, or ICARION_N_VOLTAGE_SLOTS entries
                                     (NULL = the domain's constant): the cloud's
                                     voltage waveforms, as icarion_gpu_set_waveforms
                                     on a handle of its own (analytic fields)  */
    const icarion_gas_component *mixture;
                                  /* NULL, or n_components components: the
                                     cloud's HSS gas mixture, as
                                     icarion_gpu_set_gas_mixture on a handle of
                                     its own                                   */
    const double *mixture_sigma_m2; /* [n_species][n_components]               */
    uint32_t n_components;
    uint32_t n_atoms;             /* entries of atoms_xyzr                     */
    const float *atoms_xyzr;      /* NULL, or [n_atoms][4] x, y, z, r (m, FP32):
                                     the cloud's EHSS atom table, indexed by its
                                     species' atom_offset / atom_count (geometry
                                     EHSS, FP64 clouds)                        */
    const float *field_E;         /* NULL, or [nz][ny][nx][3] E (V/m, FP32) at
                                     the nodes of a uniform grid: the cloud's
                                     field array (domain.use_field_array = 1),
                                     as icarion_gpu_upload_field_array (dtype 0)
                                     on a handle of its own; clouds passing the
                                     same pointer and grid share one device
                                     copy                                      */
    uint32_t field_n[3];          /* nodes along x, y, z (>= 2 each)           */
    uint32_t reserved1;
    double field_origin_m[3];     /* the first node                            */
    double field_spacing_m[3];    /* node spacing per axis (> 0)               */
    const icarion_waveform_desc *pressure;   /* NULL, or the cloud's gas
                                     pressure waveform, as
                                     icarion_gpu_set_pressure_waveform on a
                                     handle of its own (ABI 6)                 */
} icarion_cloud_desc;

/* Launch shape of the persistent batch kernel (per precision class). */
typedef struct icarion_batch_tuning {
    uint32_t block_size;          /* ions per tile = threads per workgroup:
                                     64, 128 or 256 (0 = tuned / default)      */
    uint32_t groups_per_cu;       /* persistent workgroups per compute unit
                                     (0 = tuned / default)                     */
} icarion_batch_tuning;

typedef struct icarion_cloud_stats {
    uint64_t ion_steps;
    uint64_t collisions;
    uint64_t alive;
    uint64_t arrived;
    double t_end_s;               /* engine time after the cloud's steps       */
    double done_ms;               /* when the cloud's last tile finished,
                                     after the first workgroup started         */
} icarion_cloud_stats;

/* Copy n_clouds clouds to the device (replaces any previous batch). */
int icarion_gpu_batch_upload(icarion_gpu *h, const icarion_cloud_desc *clouds, uint32_t n_clouds);

/* Run every cloud of the batch through all its steps in one graph launch.
 * fp64 / fp32: launch shapes of the FP64 and FP32 + Kahan classes (NULL = the
 * tuned shapes, else 256 ions x 4 workgroups per CU); EHSS-surrogate clouds
 * run in a third class at 256 ions x 2 workgroups per CU (the occupancy its
 * registers and LDS weights allow) and need icarion_gpu_surrogate_set_params
 * (or _train) on this handle first (else ICARION_ERR_STATE).  per_cloud
 * (n_clouds entries) and device_ms may be NULL. */
int icarion_gpu_batch_run(icarion_gpu *h, const icarion_batch_tuning *fp64,
                          const icarion_batch_tuning *fp32, icarion_cloud_stats *per_cloud,
                          double *device_ms);

/* Time each candidate launch shape on the resident batch with every cloud cut
 * to at most trial_steps steps (the batch state is restored after each trial)
 * and keep the fastest per precision class for icarion_gpu_batch_run.  best64 /
 * best32 and the rates (ion-steps/s) may be NULL. */
int icarion_gpu_batch_autotune(icarion_gpu *h, uint32_t trial_steps, icarion_batch_tuning *best64,
                               icarion_batch_tuning *best32, double *rate64, double *rate32);

/* Copy the batch state back into the clouds' host SoA arrays. */
int icarion_gpu_batch_download(icarion_gpu *h, const icarion_cloud_desc *clouds, uint32_t n_clouds);

/* One engine step on a host-owned ensemble: upload, one step, download.
 * The per-call shape of the reference batch hooks (IIntegrationStrategy.h:106-132,
 * ICollisionHandler.h:109-134), crossing PCIe both ways. */
int icarion_gpu_step_batch(icarion_gpu *h, icarion_ion_soa *ions, double t, double dt,
                           uint64_t step_index, icarion_step_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* ICARION_GPU_H */
