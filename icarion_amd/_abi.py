"""ctypes mirror of include/icarion_gpu.h (structs and enums).

Field order and types follow the header exactly; tests/test_abi.py checks the
sizes against offsets computed by the C compiler.
"""
import ctypes as C

ABI_VERSION = 6

OK, ERR_INVALID, ERR_HIP, ERR_UNSUPPORTED, ERR_STATE, ERR_NO_DEVICE = range(6)

INSTRUMENT_IMS, INSTRUMENT_TIMS, INSTRUMENT_QUADRUPOLE, INSTRUMENT_ORBITRAP = range(4)
INTEGRATOR_RK4, INTEGRATOR_BORIS, INTEGRATOR_RK45 = range(3)
COLLISION_NONE, COLLISION_HSS, COLLISION_EHSS, COLLISION_LANGEVIN, COLLISION_FRICTION = range(5)
PRECISION_FP64, PRECISION_FP32_KAHAN = range(2)
EHSS_GEOMETRY, EHSS_SURROGATE = range(2)
(FIELD_SCALE_CONSTANT, FIELD_SCALE_DC_AXIAL, FIELD_SCALE_DC_QUAD, FIELD_SCALE_DC_RADIAL,
 FIELD_SCALE_RF) = range(5)
MAX_FIELD_TERMS = 4
(WAVEFORM_CONSTANT, WAVEFORM_LINEAR, WAVEFORM_QUADRATIC, WAVEFORM_EXPONENTIAL, WAVEFORM_SINUSOIDAL,
 WAVEFORM_PWM, WAVEFORM_PULSED, WAVEFORM_ARBITRARY) = range(8)
INTERP_LINEAR, INTERP_STEP, INTERP_CUBIC = range(3)
FLOW_CONSTANT, FLOW_AXIAL_PARABOLIC = range(2)
WAVEFORM_LIBRARY, WAVEFORM_INLINE = range(2)
SLOT_DC_AXIAL, SLOT_DC_QUAD, SLOT_DC_RADIAL, SLOT_RF_VOLTAGE, SLOT_RF_FREQUENCY = range(5)
N_VOLTAGE_SLOTS = 5

_d = C.c_double
_i32 = C.c_int32
_u32 = C.c_uint32
_u64 = C.c_uint64


class DomainDesc(C.Structure):
    _fields_ = [
        ("instrument", _i32), ("use_field_array", _i32), ("magnetic_enabled", _i32),
        ("reserved0", _i32),
        ("origin_m", _d * 3), ("length_m", _d), ("radius_m", _d),
        ("radius_in_m", _d), ("radius_out_m", _d), ("radius_char_m", _d),
        ("rot_g2l", _d * 9), ("rot_l2g", _d * 9),
        ("dc_axial_V", _d), ("dc_quad_V", _d), ("dc_radial_V", _d),
        ("rf_voltage_V", _d), ("rf_frequency_Hz", _d), ("field_array_scale", _d),
        ("B_T", _d * 3),
        ("temperature_K", _d), ("pressure_Pa", _d), ("gas_density_m3", _d),
        ("gas_mass_kg", _d), ("gas_radius_m", _d), ("gas_polarizability_m3", _d),
        ("gas_velocity_m_s", _d * 3), ("detector_z_m", _d),
    ]


class PhysicsDesc(C.Structure):
    _fields_ = [
        ("integrator", _i32), ("collision_model", _i32), ("ou_thermal_kicks", _i32),
        ("precision", _i32), ("ehss_mode", _i32), ("collision_substeps", _i32), ("rng_seed", _u64),
    ]


class SpeciesDesc(C.Structure):
    _fields_ = [
        ("sigma_hss_m2", _d), ("sigma_ehss_m2", _d), ("ccs_m2", _d), ("K0_cm2_Vs", _d),
        ("atom_offset", _u32), ("atom_count", _u32), ("surrogate_features", C.c_float * 3),
        ("reserved0", _u32),
    ]


class TrainDesc(C.Structure):
    _fields_ = [("epochs", _u32), ("batch", _u32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", _u64)]


SURROGATE_NPARAMS = 17795


class IonSoA(C.Structure):
    _fields_ = [
        ("n", C.c_size_t),
        ("x", C.POINTER(_d)), ("y", C.POINTER(_d)), ("z", C.POINTER(_d)),
        ("vx", C.POINTER(_d)), ("vy", C.POINTER(_d)), ("vz", C.POINTER(_d)),
        ("mass_kg", C.POINTER(_d)), ("charge_C", C.POINTER(_d)),
        ("species", C.POINTER(_u32)), ("active", C.POINTER(C.c_uint8)),
        ("born", C.POINTER(C.c_uint8)), ("birth_time_s", C.POINTER(_d)),
        ("death_time_s", C.POINTER(_d)), ("arrival_time_s", C.POINTER(_d)),
        ("n_collisions", C.POINTER(_u32)),
        ("time_s", C.POINTER(_d)), ("dt_s", C.POINTER(_d)),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("ion_steps", _u64), ("collisions", _u64), ("alive", _u64), ("arrived", _u64),
        ("rejected", _u64), ("t_end_s", _d), ("device_ms", _d),
    ]


# int (*icarion_snapshot_fn)(void*, uint64_t index, uint64_t step, double t_s, size_t n,
#                           const double* pos, const double* vel, const uint32_t* species,
#                           const int32_t* domain_index, const uint8_t* flags)
SnapshotFn = C.CFUNCTYPE(C.c_int, C.c_void_p, _u64, _u64, _d, C.c_size_t, C.POINTER(_d),
                         C.POINTER(_d), C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                         C.POINTER(C.c_uint8))


class ArrivalStats(C.Structure):
    _fields_ = [("count", _u64), ("min_s", _d), ("max_s", _d), ("mean_s", _d), ("std_s", _d)]


class Rk45Desc(C.Structure):
    _fields_ = [
        ("atol", _d), ("rtol", _d), ("safety_factor", _d), ("min_step_factor", _d),
        ("max_step_factor", _d), ("max_step_increase", _d), ("max_step_decrease", _d),
        ("absolute_min_step_s", _d), ("accept_at_dt_min", _i32), ("reserved0", _i32),
    ]


def rk45_defaults(**kw):
    """RK45Strategy::AdaptiveConfig defaults (RK45Strategy.h:54-65)."""
    d = Rk45Desc(atol=1e-8, rtol=1e-6, safety_factor=0.9, min_step_factor=1e-6,
                 max_step_factor=1e3, max_step_increase=5.0, max_step_decrease=0.1,
                 absolute_min_step_s=0.0, accept_at_dt_min=1)
    for k, v in kw.items():
        setattr(d, k, v)
    return d


class BatchTuning(C.Structure):
    _fields_ = [("block_size", _u32), ("groups_per_cu", _u32)]


class CloudStats(C.Structure):
    _fields_ = [("ion_steps", _u64), ("collisions", _u64), ("alive", _u64), ("arrived", _u64),
                ("t_end_s", _d), ("done_ms", _d)]


class GpuDesc(C.Structure):
    _fields_ = [("device", _i32), ("block_size", _u32), ("steps_per_launch", _u32),
                ("reserved0", _u32)]


class GasComponent(C.Structure):
    _fields_ = [("density_m3", _d), ("mass_kg", _d), ("participates", _i32), ("reserved0", _i32)]


MAX_GAS_COMPONENTS = 4


class OracleGasMixture(C.Structure):
    _fields_ = [("n_components", _u32), ("n_species", _u32),
                ("components", C.POINTER(GasComponent)), ("sigma_m2", C.POINTER(_d)),
                ("radius_m", C.POINTER(_d))]


class OracleEhssOffline(C.Structure):
    _fields_ = [("n_species", _u32), ("reserved0", _u32), ("n_orient", C.POINTER(_u32)),
                ("mu_stride", C.POINTER(_u32)), ("sigma_eff", C.POINTER(_d)), ("mu", C.POINTER(_d))]


class OracleEhssOrientations(C.Structure):
    _fields_ = [("n_species", _u32), ("reserved0", _u32), ("n_orient", C.POINTER(_u32)),
                ("quat_wxyz", C.POINTER(_d)), ("area_m2", C.POINTER(_d))]


class OracleCollisionTables(C.Structure):
    _fields_ = [("mixture", C.POINTER(OracleGasMixture)),
                ("ehss_offline", C.POINTER(OracleEhssOffline)),
                ("ehss_orient", C.POINTER(OracleEhssOrientations))]


class WaveformDesc(C.Structure):
    _fields_ = [("type", _i32), ("interp", _i32), ("clamp", _i32), ("source", _i32),
                ("p", _d * 8), ("times_s", C.POINTER(_d)), ("values", C.POINTER(_d)),
                ("n_points", _u32), ("reserved1", _u32)]


class CloudDesc(C.Structure):
    _fields_ = [
        ("domain", DomainDesc), ("physics", PhysicsDesc),
        ("species", C.POINTER(SpeciesDesc)), ("n_species", _u32), ("reserved0", _u32),
        ("ions", C.POINTER(IonSoA)), ("t0_s", _d), ("dt_s", _d), ("n_steps", _u64),
        ("step_offset", _u64), ("waveforms", C.POINTER(C.POINTER(WaveformDesc))),
        ("mixture", C.POINTER(GasComponent)), ("mixture_sigma_m2", C.POINTER(_d)),
        ("n_components", _u32), ("n_atoms", _u32), ("atoms_xyzr", C.POINTER(C.c_float)),
        ("field_E", C.POINTER(C.c_float)), ("field_n", _u32 * 3), ("reserved1", _u32),
        ("field_origin_m", _d * 3), ("field_spacing_m", _d * 3),
        ("pressure", C.POINTER(WaveformDesc)),
    ]


def waveform(type_, p=(), clamp=1, interp=0, points=(), source=WAVEFORM_LIBRARY):
    """An icarion_waveform_desc; the returned object keeps its arrays alive."""
    import numpy as np
    w = WaveformDesc()
    w.type, w.clamp, w.interp, w.source = type_, int(clamp), interp, source
    for i, x in enumerate(p):
        w.p[i] = float(x)
    pts = np.asarray(list(points), dtype=np.float64).reshape(-1, 2)
    w._ts = np.ascontiguousarray(pts[:, 0])
    w._vs = np.ascontiguousarray(pts[:, 1])
    if len(pts):
        w.times_s = w._ts.ctypes.data_as(C.POINTER(_d))
        w.values = w._vs.ctypes.data_as(C.POINTER(_d))
    w.n_points = len(pts)
    return w


class OracleFieldTerm(C.Structure):
    _fields_ = [
        ("scale_kind", _i32), ("reserved0", _i32), ("constant_scale", _d), ("frequency_Hz", _d),
        ("phase_rad", _d), ("Ex", C.POINTER(_d)), ("Ey", C.POINTER(_d)), ("Ez", C.POINTER(_d)),
    ]


class OracleField(C.Structure):
    _fields_ = [
        ("nx", _u32), ("ny", _u32), ("nz", _u32),
        ("xs", C.POINTER(_d)), ("ys", C.POINTER(_d)), ("zs", C.POINTER(_d)),
        ("Ex", C.POINTER(_d)), ("Ey", C.POINTER(_d)), ("Ez", C.POINTER(_d)),
        ("n_terms", _u32), ("reserved0", _u32), ("terms", C.POINTER(OracleFieldTerm)),
        ("waveforms", C.POINTER(WaveformDesc) * 5),
        ("pressure", C.POINTER(WaveformDesc)),
        ("flow_model", _i32), ("reserved1", _i32), ("flow_max_velocity_m_s", _d),
    ]


class FieldTermDesc(C.Structure):
    _fields_ = [
        ("E", C.c_void_p), ("dtype", _i32), ("scale_kind", _i32), ("constant_scale", _d),
        ("frequency_Hz", _d), ("phase_rad", _d),
    ]


IDENTITY9 = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
