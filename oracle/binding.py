"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py - never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from icarion_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

RNG_PHILOX = 0
RNG_MT19937_64 = 1

_lib = None


def build(force=False):
    """Compile liboracle.so with the Makefile beside this file (g++)."""
    args = ["make", "-C", HERE, "liboracle.so"]
    if force:
        args = ["make", "-C", HERE, "-B", "liboracle.so"]
    subprocess.run(args, check=True, capture_output=True)
    return LIB


def build_ref():
    """Compile the reference-kernel harness into _ref/ (needs /root/reference)."""
    subprocess.run(["make", "-C", HERE, "ref"], check=True, capture_output=True)
    return REF_HARNESS


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    P = C.POINTER
    d = C.c_double
    lib.oracle_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    lib.oracle_uniform_from_words.argtypes = [C.c_uint32, C.c_uint32]
    lib.oracle_uniform_from_words.restype = d
    lib.oracle_philox_uniforms.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, P(d)]
    lib.oracle_field_E.argtypes = [P(_abi.DomainDesc), P(_abi.OracleField), P(d), d, P(d)]
    lib.oracle_contains.argtypes = [P(_abi.DomainDesc), P(d)]
    lib.oracle_contains.restype = C.c_int
    lib.oracle_orbitrap_r_for_z.argtypes = [d, d, d]
    lib.oracle_orbitrap_r_for_z.restype = d
    lib.oracle_interpolate_field.argtypes = [P(_abi.OracleField), P(d), P(d)]
    lib.oracle_waveform_value.argtypes = [P(_abi.WaveformDesc), d]
    lib.oracle_waveform_value.restype = d
    lib.oracle_gas_velocity.argtypes = [P(_abi.DomainDesc), P(_abi.OracleField), P(d), P(d)]
    lib.oracle_field_term_scale.argtypes = [P(_abi.DomainDesc), P(_abi.OracleFieldTerm),
                                            P(_abi.OracleField), d]
    lib.oracle_field_term_scale.restype = d
    lib.oracle_damping_gamma.argtypes = [P(_abi.DomainDesc), C.c_int, P(_abi.SpeciesDesc), d, d,
                                         P(d)]
    lib.oracle_damping_gamma.restype = d
    lib.oracle_mt_sample_neutral_velocity.argtypes = [C.c_uint64, d, d, P(d), P(d)]
    lib.oracle_mt_hss_collision.argtypes = [C.c_uint64, P(d), P(d), d, d, P(d)]
    lib.oracle_mt_ehss_collision.argtypes = [C.c_uint64, P(d), P(d), d, d, d, P(d), C.c_uint32,
                                             P(d)]
    lib.oracle_mt_ehss_oriented.argtypes = [C.c_uint64, P(d), P(d), d, d, d, P(d), C.c_uint32,
                                            P(d), d, P(d)]
    lib.oracle_quaternion_to_rotation.argtypes = [P(d), P(d)]
    lib.oracle_mt_ehss_mixture.argtypes = [C.c_uint64, P(_abi.DomainDesc), P(_abi.OracleGasMixture),
                                           C.c_uint32, P(d), C.c_uint32, d, P(d), d]
    lib.oracle_mt_ehss_mixture.restype = C.c_int
    lib.oracle_oapa_sigma.argtypes = [P(d), C.c_uint32, d]
    lib.oracle_oapa_sigma.restype = d
    lib.oracle_push_one.argtypes = [P(_abi.DomainDesc), P(_abi.PhysicsDesc), P(_abi.SpeciesDesc),
                                    P(_abi.OracleField), d, d, P(d), P(d), d, d]
    lib.oracle_run.argtypes = [P(_abi.DomainDesc), P(_abi.PhysicsDesc), P(_abi.SpeciesDesc),
                               C.c_uint32, P(d), C.c_uint32, P(_abi.OracleField), P(_abi.IonSoA),
                               d, d, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                               P(_abi.StepStats), P(_abi.OracleCollisionTables)]
    lib.oracle_run.restype = C.c_int
    lib.oracle_find_domain.argtypes = [P(_abi.DomainDesc), P(d), C.c_uint32, P(d)]
    lib.oracle_find_domain.restype = C.c_int
    lib.oracle_run_multi.argtypes = [P(_abi.DomainDesc), P(d), C.c_uint32, P(P(_abi.OracleField)),
                                     P(P(_abi.OracleGasMixture)), P(P(_abi.SpeciesDesc)), P(_abi.PhysicsDesc),
                                     P(_abi.SpeciesDesc), C.c_uint32, P(d), P(_abi.IonSoA), P(C.c_int32),
                                     d, d, C.c_uint64, C.c_uint64, C.c_int, P(_abi.StepStats)]
    lib.oracle_run_multi.restype = C.c_int
    lib.oracle_mt_ehss_offline.argtypes = [C.c_uint64, P(_abi.DomainDesc), C.c_uint32, C.c_uint32,
                                           P(d), P(d), d, P(d), d, P(C.c_uint32), P(d), P(d)]
    lib.oracle_mt_ehss_offline.restype = C.c_int
    lib.oracle_mt_hss_mixture.argtypes = [C.c_uint64, P(_abi.DomainDesc),
                                          P(_abi.OracleGasMixture), P(d), d, P(d), d, P(d)]
    lib.oracle_mt_hss_mixture.restype = C.c_int
    u32p, f32p = P(C.c_uint32), P(C.c_float)
    lib.oracle_ehss_precompute.argtypes = [f32p, u32p, u32p, C.c_uint32, d, C.c_uint32,
                                           C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, P(d),
                                           P(d)]
    lib.oracle_ehss_features.argtypes = [f32p, u32p, u32p, C.c_uint32, d, P(d), f32p]
    lib.oracle_oapa_sigma_f32.argtypes = [f32p, u32p, u32p, C.c_uint32, d, P(d)]
    lib.oracle_mlp_forward.argtypes = [f32p, f32p, C.c_size_t, f32p]
    lib.oracle_mlp_loss_grad.argtypes = [f32p, f32p, f32p, C.c_size_t, P(d)]
    lib.oracle_mlp_loss_grad.restype = d
    R = P(_abi.Rk45Desc)
    lib.oracle_rk45_error.argtypes = [R, P(d), P(d), P(d)]
    lib.oracle_rk45_error.restype = d
    lib.oracle_rk45_new_step.argtypes = [R, d, d, d, d, d]
    lib.oracle_rk45_new_step.restype = d
    lib.oracle_rk45_push_one.argtypes = [P(_abi.DomainDesc), P(_abi.PhysicsDesc), R,
                                         P(_abi.SpeciesDesc), P(_abi.OracleField), d, d, P(d),
                                         P(d), d, d, C.c_void_p, P(d), P(C.c_int)]
    lib.oracle_rk45_push_one.restype = d
    lib.oracle_run_rk45.argtypes = [P(_abi.DomainDesc), P(_abi.PhysicsDesc), R,
                                    P(_abi.SpeciesDesc), C.c_uint32, P(d), C.c_uint32,
                                    P(_abi.OracleField), P(_abi.IonSoA), C.c_void_p, d, d,
                                    C.c_uint64, C.c_uint64, C.c_int, P(_abi.StepStats),
                                    P(_abi.OracleCollisionTables)]
    lib.oracle_run_rk45.restype = C.c_int
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def vec(*xs):
    return (C.c_double * len(xs))(*xs)


class Field:
    """oracle_field built from a node-centred uniform grid E[nz, ny, nx, 3]."""

    def __init__(self, E, origin, spacing):
        E = np.asarray(E)
        nz, ny, nx = E.shape[:3]
        self.xs = origin[0] + spacing[0] * np.arange(nx, dtype=np.float64)
        self.ys = origin[1] + spacing[1] * np.arange(ny, dtype=np.float64)
        self.zs = origin[2] + spacing[2] * np.arange(nz, dtype=np.float64)
        self.Ex = np.ascontiguousarray(E[..., 0], dtype=np.float64).ravel()
        self.Ey = np.ascontiguousarray(E[..., 1], dtype=np.float64).ravel()
        self.Ez = np.ascontiguousarray(E[..., 2], dtype=np.float64).ravel()
        self.c = _abi.OracleField(nx, ny, nz, _dp(self.xs), _dp(self.ys), _dp(self.zs),
                                  _dp(self.Ex), _dp(self.Ey), _dp(self.Ez), 0, 0,
                                  C.POINTER(_abi.OracleFieldTerm)())


class FieldTerms:
    """oracle_field on a rectilinear grid (axes xs, ys, zs, possibly
    non-uniform) with terms [(E[nz, ny, nx, 3], scale_kind, constant_scale,
    frequency_Hz, phase_rad), ...] and optional voltage waveforms
    {slot: WaveformDesc}."""

    def __init__(self, axes, terms, waveforms=None):
        self.xs, self.ys, self.zs = (np.ascontiguousarray(a, dtype=np.float64) for a in axes)
        nx, ny, nz = len(self.xs), len(self.ys), len(self.zs)
        self.arrays = []
        self.terms = (_abi.OracleFieldTerm * len(terms))()
        for i, (E, kind, const, freq, phase) in enumerate(terms):
            E = np.asarray(E)
            assert E.shape[:3] == (nz, ny, nx)
            comp = [np.ascontiguousarray(E[..., c], dtype=np.float64).ravel() for c in range(3)]
            self.arrays.append(comp)
            self.terms[i] = _abi.OracleFieldTerm(kind, 0, const, freq, phase, _dp(comp[0]),
                                                 _dp(comp[1]), _dp(comp[2]))
        first = self.arrays[0]
        self.c = _abi.OracleField(nx, ny, nz, _dp(self.xs), _dp(self.ys), _dp(self.zs),
                                  _dp(first[0]), _dp(first[1]), _dp(first[2]), len(terms), 0,
                                  C.cast(self.terms, C.POINTER(_abi.OracleFieldTerm)))
        self.slots = dict(waveforms or {})
        for k, w in self.slots.items():
            self.c.waveforms[k] = C.pointer(w)


class Waveforms:
    """An oracle_field carrying only voltage waveforms (for an analytic-field
    domain): slots {slot: WaveformDesc}."""

    def __init__(self, slots):
        self.slots = dict(slots)
        self.c = _abi.OracleField()
        for k, w in self.slots.items():
            self.c.waveforms[k] = C.pointer(w)


def waveform_value(w, t):
    return load().oracle_waveform_value(C.byref(w), t)


def gas_velocity(domain, pos, flow=None):
    """gas_velocity_at of the domain at pos; flow: None or (model, max velocity)."""
    f = _with_flow(None, flow)
    out = (C.c_double * 3)()
    load().oracle_gas_velocity(C.byref(domain), C.byref(f.c) if f is not None else None, vec(*pos), out)
    return list(out)


def field_term_scale(domain, term, t, field=None):
    """scale_i(t) of one term (kind, constant, frequency, phase) for the domain;
    field (FieldTerms or Waveforms) supplies voltage waveforms."""
    lib = load()
    kind, const, freq, phase = term
    tt = _abi.OracleFieldTerm(kind, 0, const, freq, phase)
    fp = C.byref(field.c) if field is not None else C.POINTER(_abi.OracleField)()
    return lib.oracle_field_term_scale(C.byref(domain), C.byref(tt), fp, t)


def species_array(species):
    arr = (_abi.SpeciesDesc * len(species))()
    for i, s in enumerate(species):
        arr[i] = s
    return arr


class GasMixture:
    """oracle_gas_mixture: components [(density_m3, mass_kg, participates), ...],
    sigma[n_species, n_components] (m^2) and optional contact radii (m)."""

    def __init__(self, components, sigma, radii=None):
        self.comps = (_abi.GasComponent * max(len(components), 1))()
        for i, (n, m, part) in enumerate(components):
            self.comps[i] = _abi.GasComponent(n, m, int(part), 0)
        self.sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        n_species = self.sigma.shape[0] if len(components) else 0
        self.radii = None if radii is None else np.ascontiguousarray(radii, dtype=np.float64)
        self.c = _abi.OracleGasMixture(len(components), n_species,
                                       C.cast(self.comps, C.POINTER(_abi.GasComponent)),
                                       _dp(self.sigma),
                                       _dp(self.radii) if self.radii is not None else C.POINTER(C.c_double)())


class EhssOffline:
    """oracle_ehss_offline from per-species tables: sets[s] = None (online
    geometry) or (sigma_eff[n_orient], mu[n_orient, mu_stride])."""

    def __init__(self, sets):
        n = len(sets)
        self.n_orient = np.zeros(n, np.uint32)
        self.stride = np.zeros(n, np.uint32)
        sig, mus = [], []
        for s_, t in enumerate(sets):
            if t is None:
                continue
            sg, mu = np.asarray(t[0], np.float64), np.asarray(t[1], np.float64)
            self.n_orient[s_], self.stride[s_] = sg.shape[0], mu.shape[1]
            sig.append(sg.ravel())
            mus.append(mu.ravel())
        self.sigma = np.ascontiguousarray(np.concatenate(sig) if sig else np.zeros(1))
        self.mu = np.ascontiguousarray(np.concatenate(mus) if mus else np.zeros(1))
        u32 = C.POINTER(C.c_uint32)
        self.c = _abi.OracleEhssOffline(n, 0, self.n_orient.ctypes.data_as(u32),
                                        self.stride.ctypes.data_as(u32), _dp(self.sigma),
                                        _dp(self.mu))


class EhssOrientations:
    """oracle_ehss_orientations from per-species samples: sets[s] = None (no
    samples) or (quat_wxyz[n_orient, 4], area_m2[n_orient])."""

    def __init__(self, sets):
        n = len(sets)
        self.n_orient = np.zeros(n, np.uint32)
        qs, ars = [], []
        for s_, t in enumerate(sets):
            if t is None:
                continue
            q, a = np.asarray(t[0], np.float64).reshape(-1, 4), np.asarray(t[1], np.float64).ravel()
            assert q.shape[0] == a.shape[0]
            self.n_orient[s_] = q.shape[0]
            qs.append(q.ravel())
            ars.append(a)
        self.quat = np.ascontiguousarray(np.concatenate(qs) if qs else np.zeros(4))
        self.area = np.ascontiguousarray(np.concatenate(ars) if ars else np.zeros(1))
        self.c = _abi.OracleEhssOrientations(n, 0, self.n_orient.ctypes.data_as(C.POINTER(C.c_uint32)),
                                             _dp(self.quat), _dp(self.area))


def _tables_ptr(mixture=None, ehss_offline=None, ehss_orient=None):
    if mixture is None and ehss_offline is None and ehss_orient is None:
        return C.POINTER(_abi.OracleCollisionTables)()
    t = _abi.OracleCollisionTables(
        C.pointer(mixture.c) if mixture is not None else C.POINTER(_abi.OracleGasMixture)(),
        C.pointer(ehss_offline.c) if ehss_offline is not None
        else C.POINTER(_abi.OracleEhssOffline)(),
        C.pointer(ehss_orient.c) if ehss_orient is not None
        else C.POINTER(_abi.OracleEhssOrientations)())
    return C.pointer(t)


def mt_ehss_oriented(seed, vi, vn, mi, mn, rion, atoms, quat, sigma):
    """CollisionKernels::ehss_collision_with_orientation on PhysicsRng(seed)."""
    a = np.ascontiguousarray(atoms, dtype=np.float64).reshape(-1, 4)
    out = (C.c_double * 3)()
    load().oracle_mt_ehss_oriented(seed, vec(*vi), vec(*vn), mi, mn, rion, _dp(a), a.shape[0],
                                   vec(*quat), sigma, out)
    return list(out)


def mt_ehss_mixture(seed, domain, mixture, species, atoms, mass, vel, dt):
    """One EHSS collision stage over a gas mixture on PhysicsRng(seed):
    returns (collided, velocity after)."""
    a = np.ascontiguousarray(atoms, dtype=np.float64).reshape(-1, 4)
    v = vec(*vel)
    hit = load().oracle_mt_ehss_mixture(seed, C.byref(domain), C.byref(mixture.c), species, _dp(a),
                                        a.shape[0], mass, v, dt)
    return bool(hit), np.array(v[:])


def quaternion_to_rotation(quat):
    out = (C.c_double * 9)()
    load().oracle_quaternion_to_rotation(vec(*quat), out)
    return list(out)


def mt_ehss_offline(seed, domain, sigma_eff, mu, mass, vel, dt):
    """One offline-table EHSS collision stage on a fresh PhysicsRng(seed):
    returns (collided, orientation drawn, mu used, velocity after, neutral)."""
    lib = load()
    sg = np.ascontiguousarray(sigma_eff, dtype=np.float64)
    m = np.ascontiguousarray(mu, dtype=np.float64)
    v = vec(*vel)
    o = C.c_uint32()
    mu_used = C.c_double()
    vn = (C.c_double * 3)()
    hit = lib.oracle_mt_ehss_offline(seed, C.byref(domain), sg.shape[0], m.shape[1], _dp(sg),
                                     _dp(m), mass, v, dt, C.byref(o), C.byref(mu_used), vn)
    return bool(hit), o.value, mu_used.value, np.array(v[:]), np.array(vn[:])


def mt_hss_mixture(seed, domain, mixture, sigma_row, mass, vel, dt):
    """One mixture collision stage on a fresh PhysicsRng(seed): returns
    (selected component or -1, velocity after, selected neutral's velocity)."""
    lib = load()
    v = vec(*vel)
    vn = (C.c_double * 3)()
    row = np.ascontiguousarray(sigma_row, dtype=np.float64)
    sel = lib.oracle_mt_hss_mixture(seed, C.byref(domain), C.byref(mixture.c), _dp(row), mass, v,
                                    dt, vn)
    return sel, np.array(v[:]), np.array(vn[:])


def _with_pressure(field, pressure):
    """The oracle_field to pass: field (or an empty one) carrying the pressure
    waveform; field itself when pressure is None."""
    if pressure is None:
        return field
    f = copy_field(field) if field is not None else Waveforms({})
    f.c.pressure = C.pointer(pressure)
    f._pressure = pressure
    return f


def copy_field(field):
    """A shallow copy of a Field / FieldTerms / Waveforms whose oracle_field can
    be changed (the arrays stay shared)."""
    import copy
    g = copy.copy(field)
    g.c = _abi.OracleField.from_buffer_copy(field.c)
    return g


def _with_flow(field, flow):
    """field (or an empty one) carrying the gas flow (model, max velocity);
    field itself when flow is None."""
    if flow is None:
        return field
    f = copy_field(field) if field is not None else Waveforms({})
    f.c.flow_model, f.c.flow_max_velocity_m_s = int(flow[0]), float(flow[1])
    return f


def run(domain, physics, species, ens, t0, dt, n_steps, step_offset=0, atoms=None, field=None,
        rng_kind=RNG_PHILOX, threads=0, mixture=None, ehss_offline=None, ehss_orient=None, pressure=None,
        flow=None):
    """Run the oracle engine loop on ens in place; returns StepStats.
    pressure: None, or the gas pressure's WaveformDesc.  flow: None, or
    (icarion_flow_model, axial max velocity [m/s])."""
    lib = load()
    field = _with_flow(_with_pressure(field, pressure), flow)
    sp = species_array(species)
    if atoms is not None and len(atoms):
        a = np.ascontiguousarray(atoms, dtype=np.float64).reshape(-1, 4)
        ap, na = _dp(a), a.shape[0]
    else:
        a, ap, na = None, C.POINTER(C.c_double)(), 0
    fptr = C.byref(field.c) if field is not None else C.POINTER(_abi.OracleField)()
    soa = ens.soa()
    st = _abi.StepStats()
    rc = lib.oracle_run(C.byref(domain), C.byref(physics), sp, len(species), ap, na, fptr,
                        C.byref(soa), t0, dt, n_steps, step_offset, rng_kind, threads,
                        C.byref(st), _tables_ptr(mixture, ehss_offline, ehss_orient))
    if rc != 0:
        raise RuntimeError(f"oracle_run failed ({rc})")
    return st


def _domains(domains, end_apertures):
    arr = (_abi.DomainDesc * len(domains))(*domains)
    ap = None if end_apertures is None else np.ascontiguousarray(end_apertures, dtype=np.float64)
    app = _dp(ap) if ap is not None else C.POINTER(C.c_double)()
    return arr, ap, app


def find_domain(domains, pos, end_apertures=None):
    """DomainManager::find_domain_index of pos over the domains (-1: none)."""
    arr, ap, app = _domains(domains, end_apertures)
    return load().oracle_find_domain(arr, app, len(domains), vec(*pos))


def run_multi(domains, physics, species, ens, domain_index, t0, dt, n_steps, step_offset=0,
              end_apertures=None, threads=0, waveforms=None, atoms=None, mixtures=None,
              domain_species=None, fields=None, pressures=None, flows=None):
    """The engine loop over several domains on ens and domain_index (int32,
    in/out) in place; returns StepStats.  waveforms: None, or per domain None
    or {slot: WaveformDesc}.  domain_species: None, or per domain None or a
    species list (that domain's rows of the species table).  fields: None, or
    per domain None or a Field (the field array of a domain with
    use_field_array = 1).  pressures: None, or per domain None or the
    domain's pressure WaveformDesc.  flows: None, or per domain None or
    (icarion_flow_model, axial max velocity)."""
    lib = load()
    arr, ap, app = _domains(domains, end_apertures)
    sp = species_array(species)
    soa = ens.soa()
    st = _abi.StepStats()
    assert domain_index.dtype == np.int32 and domain_index.flags.c_contiguous
    fptr = C.POINTER(C.POINTER(_abi.OracleField))()
    keep = []
    if waveforms is not None or fields is not None or pressures is not None or flows is not None:
        fl = (C.POINTER(_abi.OracleField) * len(domains))()
        for k in range(len(domains)):
            slots = waveforms[k] if waveforms is not None else None
            fld = fields[k] if fields is not None else None
            pw = pressures[k] if pressures is not None else None
            fw = flows[k] if flows is not None else None
            assert not (slots and fld is not None), "a domain takes a field array or waveforms"
            if fld is None and (slots or pw is not None or fw is not None):
                fld = Waveforms(slots or {})
            if fld is not None:
                fld = _with_flow(_with_pressure(fld, pw), fw)
                keep.append(fld)
                fl[k] = C.pointer(fld.c)
        fptr = fl
    mptr = C.POINTER(C.POINTER(_abi.OracleGasMixture))()
    if mixtures is not None:
        ml = (C.POINTER(_abi.OracleGasMixture) * len(domains))()
        for k, mx in enumerate(mixtures):
            if mx:
                gm = GasMixture(*mx)
                keep.append(gm)
                ml[k] = C.pointer(gm.c)
        mptr = ml
    sptr = C.POINTER(C.POINTER(_abi.SpeciesDesc))()
    if domain_species is not None:
        sl = (C.POINTER(_abi.SpeciesDesc) * len(domains))()
        for k, rows in enumerate(domain_species):
            if rows:
                assert len(rows) == len(species)
                ra = species_array(rows)
                keep.append(ra)
                sl[k] = C.cast(ra, C.POINTER(_abi.SpeciesDesc))
        sptr = sl
    at = None
    if atoms is not None:
        at = np.ascontiguousarray(atoms, dtype=np.float64).reshape(-1)
    aptr = at.ctypes.data_as(C.POINTER(C.c_double)) if at is not None else C.POINTER(C.c_double)()
    rc = lib.oracle_run_multi(arr, app, len(domains), fptr, mptr, sptr, C.byref(physics), sp, len(species), aptr, C.byref(soa),
                              domain_index.ctypes.data_as(C.POINTER(C.c_int32)), t0, dt, n_steps,
                              step_offset, threads, C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle_run_multi failed ({rc})")
    return st


def field_E(domain, pos, t, field=None):
    lib = load()
    out = (C.c_double * 3)()
    fptr = C.byref(field.c) if field is not None else C.POINTER(_abi.OracleField)()
    lib.oracle_field_E(C.byref(domain), fptr, vec(*pos), t, out)
    return np.array(out[:])


def philox(ctr, key):
    lib = load()
    out = (C.c_uint32 * 4)()
    lib.oracle_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    return list(out)


def philox_uniforms(seed, ion, step, n):
    lib = load()
    out = (C.c_double * n)()
    lib.oracle_philox_uniforms(seed, ion, step, n, out)
    return np.array(out[:])


def push_one(domain, physics, sp, mass, charge, pos, vel, t, dt, field=None):
    lib = load()
    p = vec(*pos)
    v = vec(*vel)
    fptr = C.byref(field.c) if field is not None else C.POINTER(_abi.OracleField)()
    lib.oracle_push_one(C.byref(domain), C.byref(physics), C.byref(sp), fptr, mass, charge, p, v,
                        t, dt)
    return np.array(p[:]), np.array(v[:])


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def cluster_offsets(counts):
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    offs = np.zeros(len(counts), np.uint32)
    offs[1:] = np.cumsum(counts[:-1])
    return offs, counts


def ehss_precompute(atoms, counts, r_gas, n_orient, n_sigma, n_mu, max_tries=10000, seed=0):
    lib = load()
    atoms = np.ascontiguousarray(atoms, dtype=np.float32)
    offs, counts = cluster_offsets(counts)
    n = len(counts)
    sig = np.zeros((n, n_orient))
    mu = np.zeros((n, n_orient, max(n_mu, 1)))
    lib.oracle_ehss_precompute(_fp(atoms), _up(offs), _up(counts), n, r_gas, n_orient, n_sigma,
                               n_mu, max_tries, seed, _dp(sig), _dp(mu))
    return sig, mu[..., :n_mu]


def oapa_sigma(atoms, counts, r_gas):
    lib = load()
    atoms = np.ascontiguousarray(atoms, dtype=np.float32)
    offs, counts = cluster_offsets(counts)
    sig = np.zeros(len(counts))
    lib.oracle_oapa_sigma_f32(_fp(atoms), _up(offs), _up(counts), len(counts), r_gas, _dp(sig))
    return sig


def ehss_features(atoms, counts, r_gas, sigma):
    lib = load()
    atoms = np.ascontiguousarray(atoms, dtype=np.float32)
    offs, counts = cluster_offsets(counts)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    f = np.zeros((len(counts), 3), np.float32)
    lib.oracle_ehss_features(_fp(atoms), _up(offs), _up(counts), len(counts), r_gas, _dp(sigma),
                             _fp(f))
    return f


def mlp_forward(params, X):
    lib = load()
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.zeros((X.shape[0], 3), np.float32)
    lib.oracle_mlp_forward(_fp(params), _fp(X), X.shape[0], _fp(Y))
    return Y


def mlp_loss_grad(params, X, T):
    lib = load()
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    g = np.zeros(params.size)
    loss = lib.oracle_mlp_loss_grad(_fp(params), _fp(X), _fp(T), X.shape[0], _dp(g))
    return loss, g


# oracle_rk45_ion: t, dt, last_error, k1a[3], fsal, reserved
RK45_ION = np.dtype([("t", "f8"), ("dt", "f8"), ("last_error", "f8"), ("k1a", "f8", 3),
                     ("fsal", "i4"), ("reserved", "i4")])


def rk45_state(ens):
    """Fresh per-ion RK45 state from an ensemble (time_s NaN and dt_s <= 0 are
    filled in by the run from t0 and dt)."""
    st = np.zeros(ens.n, RK45_ION)
    st["t"] = ens.time_s
    st["dt"] = ens.dt_s
    st["last_error"] = 1.0
    return st


def rk45_error(cfg, y4, y5, y0):
    lib = load()
    return lib.oracle_rk45_error(C.byref(cfg), vec(*y4), vec(*y5), vec(*y0))


def rk45_new_step(cfg, dt, err, dt_min, dt_max, prev_err):
    return load().oracle_rk45_new_step(C.byref(cfg), dt, err, dt_min, dt_max, prev_err)


def rk45_push_one(domain, physics, cfg, sp, mass, charge, pos, vel, t, dt, state=None,
                  field=None):
    """One RK45Strategy::step; returns (pos, vel, dt_used, dt_next, rejected, state)."""
    lib = load()
    st = np.zeros(1, RK45_ION) if state is None else np.array(state, RK45_ION).reshape(1)
    if state is None:
        st["last_error"] = 1.0
    p, v = vec(*pos), vec(*vel)
    dn = C.c_double()
    rej = C.c_int()
    fptr = C.byref(field.c) if field is not None else C.POINTER(_abi.OracleField)()
    used = lib.oracle_rk45_push_one(C.byref(domain), C.byref(physics), C.byref(cfg),
                                    C.byref(sp), fptr, mass, charge, p, v, t, dt,
                                    st.ctypes.data, C.byref(dn), C.byref(rej))
    return np.array(p[:]), np.array(v[:]), used, dn.value, rej.value, st[0]


def run_rk45(domain, physics, cfg, species, ens, t0, dt, n_steps, step_offset=0, atoms=None,
             field=None, state=None, threads=0, mixture=None, ehss_offline=None):
    """The RK45 engine loop on ens in place; returns (StepStats, state).  Pass
    the returned state to the next call to continue a run."""
    lib = load()
    sp = species_array(species)
    if atoms is not None and len(atoms):
        a = np.ascontiguousarray(atoms, dtype=np.float64).reshape(-1, 4)
        ap, na = _dp(a), a.shape[0]
    else:
        a, ap, na = None, C.POINTER(C.c_double)(), 0
    fptr = C.byref(field.c) if field is not None else C.POINTER(_abi.OracleField)()
    st = rk45_state(ens) if state is None else state
    soa = ens.soa()
    stats = _abi.StepStats()
    rc = lib.oracle_run_rk45(C.byref(domain), C.byref(physics), C.byref(cfg), sp, len(species),
                             ap, na, fptr, C.byref(soa), st.ctypes.data, t0, dt, n_steps,
                             step_offset, threads, C.byref(stats),
                             _tables_ptr(mixture, ehss_offline))
    if rc != 0:
        raise RuntimeError(f"oracle_run_rk45 failed ({rc})")
    return stats, st
