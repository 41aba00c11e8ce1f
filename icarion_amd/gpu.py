"""ctypes binding of the C ABI (include/icarion_gpu.h) of libicarion_gpu.so.

Loading fails loudly when the library is missing: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

from . import _abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libicarion_gpu.so")

_lib = None

EXPORTED = (
    "icarion_gpu_abi_version", "icarion_gpu_device_count", "icarion_gpu_last_error",
    "icarion_gpu_create", "icarion_gpu_destroy", "icarion_gpu_set_domain",
    "icarion_gpu_set_species", "icarion_gpu_upload_field_array", "icarion_gpu_upload_state",
    "icarion_gpu_step", "icarion_gpu_download_state", "icarion_gpu_step_batch",
    "icarion_gpu_synchronize", "icarion_gpu_synth_clusters", "icarion_gpu_ehss_precompute",
    "icarion_gpu_surrogate_set_params", "icarion_gpu_surrogate_get_params",
    "icarion_gpu_surrogate_train", "icarion_gpu_surrogate_grad", "icarion_gpu_surrogate_infer",
    "icarion_gpu_set_rk45", "icarion_gpu_batch_upload", "icarion_gpu_batch_run",
    "icarion_gpu_batch_autotune", "icarion_gpu_batch_download", "icarion_gpu_upload_field_terms",
    "icarion_gpu_set_gas_mixture", "icarion_gpu_set_ehss_offline", "icarion_gpu_set_waveforms",
    "icarion_gpu_step_snapshots", "icarion_gpu_set_domains", "icarion_gpu_download_domains",
    "icarion_gpu_set_ehss_orientations", "icarion_gpu_set_gas_radii", "icarion_gpu_arrival_stats",
    "icarion_gpu_arrival_histogram", "icarion_gpu_set_domain_waveforms", "icarion_gpu_set_domain_gas_mixture",
    "icarion_gpu_step_noalloc", "icarion_gpu_set_domain_species", "icarion_gpu_upload_domain_field_array",
    "icarion_gpu_set_pressure_waveform", "icarion_gpu_set_domain_pressure_waveform", "icarion_gpu_set_gas_flow",
    "icarion_gpu_set_domain_gas_flow",
)

# icarion_gpu_step_noalloc's per-call step limit (ICARION_NOALLOC_MAX_STEPS)
NOALLOC_MAX_STEPS = 16384


class IcarionError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"icarion_gpu error {code}: {msg}")
        self.code = code


def load(path=None):
    """Load libicarion_gpu.so (built by icarion_amd.build); raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} not found: build it with `python -m icarion_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(p)
    P = C.POINTER
    lib.icarion_gpu_abi_version.restype = C.c_int
    lib.icarion_gpu_device_count.restype = C.c_int
    lib.icarion_gpu_last_error.argtypes = [C.c_char_p, C.c_size_t]
    lib.icarion_gpu_last_error.restype = C.c_size_t
    lib.icarion_gpu_create.argtypes = [P(_abi.GpuDesc), P(C.c_void_p)]
    lib.icarion_gpu_destroy.argtypes = [C.c_void_p]
    lib.icarion_gpu_set_domain.argtypes = [C.c_void_p, P(_abi.DomainDesc), P(_abi.PhysicsDesc)]
    lib.icarion_gpu_set_rk45.argtypes = [C.c_void_p, P(_abi.Rk45Desc)]
    lib.icarion_gpu_batch_upload.argtypes = [C.c_void_p, P(_abi.CloudDesc), C.c_uint32]
    lib.icarion_gpu_batch_run.argtypes = [C.c_void_p, P(_abi.BatchTuning), P(_abi.BatchTuning),
                                          P(_abi.CloudStats), P(C.c_double)]
    lib.icarion_gpu_batch_autotune.argtypes = [C.c_void_p, C.c_uint32, P(_abi.BatchTuning),
                                               P(_abi.BatchTuning), P(C.c_double), P(C.c_double)]
    lib.icarion_gpu_batch_download.argtypes = [C.c_void_p, P(_abi.CloudDesc), C.c_uint32]
    lib.icarion_gpu_set_species.argtypes = [C.c_void_p, P(_abi.SpeciesDesc), C.c_uint32,
                                            C.c_void_p, C.c_uint32, C.c_int32]
    f32p, f64p = P(C.c_float), P(C.c_double)
    lib.icarion_gpu_synth_clusters.argtypes = [C.c_void_p, C.c_uint32, P(C.c_uint32), C.c_double,
                                               C.c_double, C.c_uint64, f32p, f64p, f32p]
    lib.icarion_gpu_ehss_precompute.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                                C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                                f64p, f64p]
    lib.icarion_gpu_surrogate_set_params.argtypes = [C.c_void_p, f32p, C.c_uint32]
    lib.icarion_gpu_surrogate_get_params.argtypes = [C.c_void_p, f32p, C.c_uint32]
    lib.icarion_gpu_surrogate_train.argtypes = [C.c_void_p, f32p, f32p, C.c_size_t,
                                                P(_abi.TrainDesc), f32p]
    lib.icarion_gpu_surrogate_grad.argtypes = [C.c_void_p, f32p, f32p, C.c_size_t, f32p, f32p]
    lib.icarion_gpu_surrogate_infer.argtypes = [C.c_void_p, f32p, C.c_size_t, f32p]
    lib.icarion_gpu_upload_field_array.argtypes = [
        C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
        C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_int32]
    lib.icarion_gpu_upload_state.argtypes = [C.c_void_p, P(_abi.IonSoA)]
    lib.icarion_gpu_step.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_uint64, C.c_uint64,
                                     P(_abi.StepStats)]
    lib.icarion_gpu_step_noalloc.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_uint64,
                                             C.c_uint64, P(_abi.StepStats)]
    lib.icarion_gpu_step_snapshots.argtypes = [
        C.c_void_p, C.c_double, C.c_double, C.c_uint64, C.c_uint64, C.c_double, C.c_int,
        P(C.c_double), _abi.SnapshotFn, C.c_void_p, P(_abi.StepStats)]
    lib.icarion_gpu_download_state.argtypes = [C.c_void_p, P(_abi.IonSoA)]
    lib.icarion_gpu_set_domains.argtypes = [C.c_void_p, P(_abi.DomainDesc), P(C.c_double), C.c_uint32,
                                            P(_abi.PhysicsDesc)]
    lib.icarion_gpu_download_domains.argtypes = [C.c_void_p, P(C.c_int32)]
    lib.icarion_gpu_set_ehss_orientations.argtypes = [C.c_void_p, P(C.c_uint32), P(C.c_double),
                                                      P(C.c_double), C.c_uint32]
    lib.icarion_gpu_set_gas_radii.argtypes = [C.c_void_p, P(C.c_double), C.c_uint32]
    lib.icarion_gpu_arrival_stats.argtypes = [C.c_void_p, P(_abi.ArrivalStats)]
    lib.icarion_gpu_arrival_histogram.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_uint32,
                                                  P(C.c_uint64)]
    lib.icarion_gpu_synchronize.argtypes = [C.c_void_p]
    lib.icarion_gpu_step_batch.argtypes = [C.c_void_p, P(_abi.IonSoA), C.c_double, C.c_double,
                                           C.c_uint64, P(_abi.StepStats)]
    for name in EXPORTED:
        if name not in ("icarion_gpu_abi_version", "icarion_gpu_device_count",
                        "icarion_gpu_last_error"):
            getattr(lib, name).restype = C.c_int
    if lib.icarion_gpu_abi_version() != _abi.ABI_VERSION:
        raise RuntimeError("libicarion_gpu.so ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


def last_error(lib=None):
    lib = lib or load()
    buf = C.create_string_buffer(4096)
    lib.icarion_gpu_last_error(buf, 4096)
    return buf.value.decode(errors="replace")


def _check(rc, lib):
    if rc != _abi.OK:
        raise IcarionError(rc, last_error(lib))


def species_array(species_list):
    arr = (_abi.SpeciesDesc * len(species_list))()
    for i, s in enumerate(species_list):
        arr[i] = s
    return arr


class Handle:
    """One device-resident ion cloud in one domain (an icarion_gpu handle)."""

    def __init__(self, device=0, block_size=0, steps_per_launch=0):
        self.lib = load()
        desc = _abi.GpuDesc(device=device, block_size=block_size,
                            steps_per_launch=steps_per_launch)
        h = C.c_void_p()
        _check(self.lib.icarion_gpu_create(C.byref(desc), C.byref(h)), self.lib)
        self.h = h
        self._keep = []
        self._batch_n = 0

    def close(self):
        if self.h:
            self.lib.icarion_gpu_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_domain(self, domain, physics):
        _check(self.lib.icarion_gpu_set_domain(self.h, C.byref(domain), C.byref(physics)),
               self.lib)

    def set_domains(self, domains, physics, end_apertures=None):
        """Several domains (a list of DomainDesc) with optional end apertures (m)."""
        arr = (_abi.DomainDesc * len(domains))(*domains)
        ap = None
        if end_apertures is not None:
            ap = np.ascontiguousarray(end_apertures, dtype=np.float64)
            assert len(ap) == len(domains)
        app = ap.ctypes.data_as(C.POINTER(C.c_double)) if ap is not None else C.POINTER(C.c_double)()
        _check(self.lib.icarion_gpu_set_domains(self.h, arr, app, len(domains), C.byref(physics)),
               self.lib)

    def download_domains(self, n):
        """Every resident ion's domain index (n = the resident count)."""
        out = np.empty(n, dtype=np.int32)
        _check(self.lib.icarion_gpu_download_domains(
            self.h, out.ctypes.data_as(C.POINTER(C.c_int32))), self.lib)
        return out

    def set_rk45(self, desc):
        """RK45 step control (an _abi.Rk45Desc, e.g. _abi.rk45_defaults(atol=...))."""
        _check(self.lib.icarion_gpu_set_rk45(self.h, C.byref(desc)), self.lib)

    # ---- batched clouds ------------------------------------------------------
    def _cloud_array(self, clouds):
        """clouds: list of dicts with domain, physics, species (list), ensemble,
        t0, dt, n_steps and optional step_offset, waveforms ({slot: desc}),
        mixture ((components, sigma) as set_gas_mixture takes them), atoms
        ((n, 4) x, y, z, r for geometry EHSS, held in FP32) and field ((E, origin,
        spacing) as upload_field_array takes them, E held in FP32) and pressure
        (a WaveformDesc, as set_pressure_waveform)."""
        arr = (_abi.CloudDesc * len(clouds))()
        keep = []
        fields = {}   # one FP32 copy per field array, so that clouds share it
        for k, c in enumerate(clouds):
            sp = species_array(c["species"])
            soa = c["ensemble"].soa()
            keep += [sp, soa, c["ensemble"]]
            arr[k].domain = c["domain"]
            arr[k].physics = c["physics"]
            arr[k].species = C.cast(sp, C.POINTER(_abi.SpeciesDesc))
            arr[k].n_species = len(c["species"])
            arr[k].ions = C.pointer(soa)
            arr[k].t0_s = c.get("t0", 0.0)
            arr[k].dt_s = c["dt"]
            arr[k].n_steps = c["n_steps"]
            arr[k].step_offset = c.get("step_offset", 0)
            if c.get("pressure") is not None:
                arr[k].pressure = C.pointer(c["pressure"])
                keep.append(c["pressure"])
            if c.get("waveforms"):
                wa = (C.POINTER(_abi.WaveformDesc) * _abi.N_VOLTAGE_SLOTS)()
                for slot, w in c["waveforms"].items():
                    wa[slot] = C.pointer(w)
                keep += [wa, dict(c["waveforms"])]
                arr[k].waveforms = C.cast(wa, C.POINTER(C.POINTER(_abi.WaveformDesc)))
            if c.get("mixture"):
                comps, sigma = c["mixture"]
                ga = (_abi.GasComponent * len(comps))()
                for i, (n, m, part) in enumerate(comps):
                    ga[i] = _abi.GasComponent(n, m, int(part), 0)
                sg = np.ascontiguousarray(sigma, dtype=np.float64)
                keep += [ga, sg]
                arr[k].mixture = C.cast(ga, C.POINTER(_abi.GasComponent))
                arr[k].mixture_sigma_m2 = sg.ctypes.data_as(C.POINTER(C.c_double))
                arr[k].n_components = len(comps)
            if c.get("atoms") is not None and len(c["atoms"]):
                at = np.ascontiguousarray(c["atoms"], dtype=np.float32).reshape(-1, 4)
                keep.append(at)
                arr[k].atoms_xyzr = at.ctypes.data_as(C.POINTER(C.c_float))
                arr[k].n_atoms = at.shape[0]
            if c.get("field") is not None:
                E, origin, spacing = c["field"]
                key = id(E)
                if key not in fields:
                    fields[key] = np.ascontiguousarray(E, dtype=np.float32)
                Ef = fields[key]
                if Ef.ndim != 4 or Ef.shape[3] != 3:
                    raise ValueError("field E must be [nz][ny][nx][3]")
                keep.append(Ef)
                arr[k].field_E = Ef.ctypes.data_as(C.POINTER(C.c_float))
                nz, ny, nx = Ef.shape[:3]
                for a, v in enumerate((nx, ny, nz)):
                    arr[k].field_n[a] = v
                    arr[k].field_origin_m[a] = float(origin[a])
                    arr[k].field_spacing_m[a] = float(spacing[a])
        return arr, keep

    def batch_upload(self, clouds):
        arr, keep = self._cloud_array(clouds)
        _check(self.lib.icarion_gpu_batch_upload(self.h, arr, len(clouds)), self.lib)
        self._batch_n = len(clouds)

    def batch_run(self, fp64=None, fp32=None):
        """Returns (list of CloudStats, device_ms).  fp64 / fp32: (block_size,
        groups_per_cu) launch shapes or None for the tuned ones."""
        def tun(t):
            return C.byref(_abi.BatchTuning(*t)) if t is not None else C.POINTER(_abi.BatchTuning)()
        stats = (_abi.CloudStats * self._batch_n)()
        ms = C.c_double()
        _check(self.lib.icarion_gpu_batch_run(self.h, tun(fp64), tun(fp32), stats, C.byref(ms)),
               self.lib)
        return list(stats), ms.value

    def batch_autotune(self, trial_steps=200):
        """Returns ((block, groups) fp64, (block, groups) fp32, rate64, rate32)."""
        b64, b32 = _abi.BatchTuning(), _abi.BatchTuning()
        r64, r32 = C.c_double(), C.c_double()
        _check(self.lib.icarion_gpu_batch_autotune(self.h, trial_steps, C.byref(b64), C.byref(b32),
                                                   C.byref(r64), C.byref(r32)), self.lib)
        return ((b64.block_size, b64.groups_per_cu), (b32.block_size, b32.groups_per_cu),
                r64.value, r32.value)

    def batch_download(self, clouds):
        arr, keep = self._cloud_array(clouds)
        _check(self.lib.icarion_gpu_batch_download(self.h, arr, len(clouds)), self.lib)

    def set_species(self, species_list, atoms_xyzr=None):
        """atoms_xyzr: (n_atoms, 4) float32 or float64 array (x, y, z, r in metres)."""
        arr = species_array(species_list)
        dtype = 1
        if atoms_xyzr is None or len(atoms_xyzr) == 0:
            atoms, n_atoms = None, 0
        else:
            a = np.asarray(atoms_xyzr)
            dtype = 0 if a.dtype == np.float32 else 1
            a = np.ascontiguousarray(a, dtype=np.float32 if dtype == 0 else np.float64)
            a = a.reshape(-1, 4)
            self._keep.append(a)
            atoms, n_atoms = a.ctypes.data_as(C.c_void_p), a.shape[0]
        _check(self.lib.icarion_gpu_set_species(self.h, arr, len(species_list), atoms, n_atoms,
                                                dtype), self.lib)

    def synth_clusters(self, counts, r_min=1.5e-10, r_max=2.0e-10, seed=0, want_atoms=False):
        """Random rigid clusters installed as the species table; returns
        (atoms or None, sigma_oapa, features)."""
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        n = len(counts)
        sig = np.zeros(n, np.float64)
        feat = np.zeros((n, 3), np.float32)
        atoms = np.zeros((int(counts.sum()), 4), np.float32) if want_atoms else None
        f32p = C.POINTER(C.c_float)
        rc = self.lib.icarion_gpu_synth_clusters(
            self.h, n, counts.ctypes.data_as(C.POINTER(C.c_uint32)), r_min, r_max, seed,
            atoms.ctypes.data_as(f32p) if atoms is not None else f32p(),
            sig.ctypes.data_as(C.POINTER(C.c_double)), feat.ctypes.data_as(f32p))
        _check(rc, self.lib)
        return atoms, sig, feat

    def ehss_precompute(self, cluster_begin, n_clusters, n_orient, n_sigma, n_mu, max_tries=10000,
                        seed=0):
        sig = np.zeros((n_clusters, n_orient), np.float64)
        mu = np.zeros((n_clusters, n_orient, max(n_mu, 1)), np.float64)
        d = C.POINTER(C.c_double)
        rc = self.lib.icarion_gpu_ehss_precompute(self.h, cluster_begin, n_clusters, n_orient,
                                                  n_sigma, n_mu, max_tries, seed,
                                                  sig.ctypes.data_as(d), mu.ctypes.data_as(d))
        _check(rc, self.lib)
        return sig, mu[..., :n_mu]

    def surrogate_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(self.lib.icarion_gpu_surrogate_set_params(
            self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size), self.lib)

    def surrogate_get_params(self):
        p = np.zeros(_abi.SURROGATE_NPARAMS, np.float32)
        _check(self.lib.icarion_gpu_surrogate_get_params(
            self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size), self.lib)
        return p

    def surrogate_train(self, X, T, epochs=10, batch=4096, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        d = _abi.TrainDesc(epochs=epochs, batch=batch, lr=lr, beta1=0.9, beta2=0.999, eps=1e-8,
                           seed=seed)
        losses = np.zeros(epochs, np.float32)
        f32p = C.POINTER(C.c_float)
        _check(self.lib.icarion_gpu_surrogate_train(self.h, X.ctypes.data_as(f32p),
                                                    T.ctypes.data_as(f32p), X.shape[0],
                                                    C.byref(d), losses.ctypes.data_as(f32p)),
               self.lib)
        return losses

    def surrogate_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(_abi.SURROGATE_NPARAMS, np.float32)
        loss = C.c_float()
        f32p = C.POINTER(C.c_float)
        _check(self.lib.icarion_gpu_surrogate_grad(self.h, X.ctypes.data_as(f32p),
                                                   T.ctypes.data_as(f32p), X.shape[0],
                                                   g.ctypes.data_as(f32p), C.byref(loss)), self.lib)
        return float(loss.value), g

    def surrogate_infer(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 3), np.float32)
        f32p = C.POINTER(C.c_float)
        _check(self.lib.icarion_gpu_surrogate_infer(self.h, X.ctypes.data_as(f32p), X.shape[0],
                                                    Y.ctypes.data_as(f32p)), self.lib)
        return Y

    def upload_field_array(self, E, origin, spacing):
        """E: array (nz, ny, nx, 3) or (N, 3) in x-fastest order, float32 or float64."""
        E = np.ascontiguousarray(E)
        nz, ny, nx = E.shape[:3]
        dtype = 0 if E.dtype == np.float32 else 1
        if dtype == 1:
            E = E.astype(np.float64, copy=False)
        rc = self.lib.icarion_gpu_upload_field_array(
            self.h, nx, ny, nz, origin[0], spacing[0], origin[1], spacing[1], origin[2],
            spacing[2], E.ctypes.data_as(C.c_void_p), dtype)
        _check(rc, self.lib)

    def upload_domain_field_array(self, domain, E, origin, spacing):
        """One domain's field array on a multi-domain handle (its descriptor has
        use_field_array = 1); E, origin, spacing as upload_field_array."""
        E = np.ascontiguousarray(E)
        nz, ny, nx = E.shape[:3]
        dtype = 0 if E.dtype == np.float32 else 1
        if dtype == 1:
            E = E.astype(np.float64, copy=False)
        self._keep.append(E)
        d = C.c_double
        rc = self.lib.icarion_gpu_upload_domain_field_array(
            self.h, C.c_uint32(domain), C.c_uint32(nx), C.c_uint32(ny), C.c_uint32(nz), d(origin[0]),
            d(spacing[0]), d(origin[1]), d(spacing[1]), d(origin[2]), d(spacing[2]),
            E.ctypes.data_as(C.c_void_p), C.c_int32(dtype))
        _check(rc, self.lib)

    def set_gas_mixture(self, components, sigma):
        """components: [(density_m3, mass_kg, participates), ...];
        sigma: [n_species, n_components] cross sections (m^2).  An empty list
        returns to the single-gas path."""
        arr = (_abi.GasComponent * max(len(components), 1))()
        for i, (n, m, part) in enumerate(components):
            arr[i] = _abi.GasComponent(n, m, int(part), 0)
        sig = np.ascontiguousarray(sigma if len(components) else np.zeros((0, 0)), dtype=np.float64)
        n_species = sig.shape[0] if len(components) else 0
        _check(self.lib.icarion_gpu_set_gas_mixture(
            self.h, arr, len(components), sig.ctypes.data_as(C.POINTER(C.c_double)), n_species),
            self.lib)

    def set_waveforms(self, slots):
        """slots: {icarion_voltage_slot: _abi.waveform(...)}; {} or None clears."""
        if not slots:
            _check(self.lib.icarion_gpu_set_waveforms(self.h, None), self.lib)
            return
        arr = (C.POINTER(_abi.WaveformDesc) * _abi.N_VOLTAGE_SLOTS)()
        for k, w in slots.items():
            arr[k] = C.pointer(w)
        self._keep.append(dict(slots))
        _check(self.lib.icarion_gpu_set_waveforms(self.h, arr), self.lib)

    def set_pressure_waveform(self, w):
        """The gas pressure as a waveform (_abi.waveform(...)), evaluated at
        every step start; None clears."""
        if w is not None:
            self._keep.append(w)
        _check(self.lib.icarion_gpu_set_pressure_waveform(self.h, None if w is None else C.pointer(w)), self.lib)

    def set_gas_flow(self, model, axial_max_velocity_m_s=0.0):
        """The gas flow profile (_abi.FLOW_CONSTANT clears, _abi.FLOW_AXIAL_PARABOLIC
        with its centre-line speed)."""
        _check(self.lib.icarion_gpu_set_gas_flow(self.h, C.c_int32(model), C.c_double(axial_max_velocity_m_s)),
               self.lib)

    def set_domain_gas_flow(self, domain, model, axial_max_velocity_m_s=0.0):
        """One domain's gas flow on a multi-domain handle (as set_gas_flow)."""
        _check(self.lib.icarion_gpu_set_domain_gas_flow(self.h, C.c_uint32(domain), C.c_int32(model),
                                                        C.c_double(axial_max_velocity_m_s)), self.lib)

    def set_domain_pressure_waveform(self, domain, w):
        """One domain's pressure waveform on a multi-domain handle; None clears."""
        if w is not None:
            self._keep.append(w)
        _check(self.lib.icarion_gpu_set_domain_pressure_waveform(
            self.h, C.c_uint32(domain), None if w is None else C.pointer(w)), self.lib)

    def set_domain_gas_mixture(self, domain, components, sigma):
        """One domain's HSS gas mixture on a multi-domain handle, arguments as
        set_gas_mixture; an empty list clears that domain's."""
        arr = (_abi.GasComponent * max(len(components), 1))()
        for i, (n, m, part) in enumerate(components):
            arr[i] = _abi.GasComponent(n, m, int(part), 0)
        sig = np.ascontiguousarray(sigma if len(components) else np.zeros((0, 0)), dtype=np.float64)
        n_species = sig.shape[0] if len(components) else 0
        _check(self.lib.icarion_gpu_set_domain_gas_mixture(
            self.h, C.c_uint32(domain), arr, len(components), sig.ctypes.data_as(C.POINTER(C.c_double)),
            n_species), self.lib)

    def set_domain_species(self, domain, species_list):
        """One domain's species table on a multi-domain handle (the handle's
        species, in order, with that domain's cross sections, CCS and K0);
        an empty list clears that domain's."""
        if not species_list:
            _check(self.lib.icarion_gpu_set_domain_species(self.h, C.c_uint32(domain), None, 0), self.lib)
            return
        arr = species_array(species_list)
        _check(self.lib.icarion_gpu_set_domain_species(self.h, C.c_uint32(domain), arr, len(species_list)),
               self.lib)

    def set_domain_waveforms(self, domain, slots):
        """One domain's waveforms on a multi-domain handle: slots as in
        set_waveforms; {} or None clears that domain's."""
        arr = None
        if slots:
            arr = (C.POINTER(_abi.WaveformDesc) * _abi.N_VOLTAGE_SLOTS)()
            for k, w in slots.items():
                arr[k] = C.pointer(w)
            self._keep.append(dict(slots))
        _check(self.lib.icarion_gpu_set_domain_waveforms(self.h, C.c_uint32(domain), arr), self.lib)

    def set_ehss_offline(self, sets):
        """sets[s]: None (species s keeps the online geometry) or
        (sigma_eff[n_orient], mu[n_orient, mu_stride]); [] clears."""
        n = len(sets)
        n_orient = np.zeros(max(n, 1), np.uint32)
        stride = np.zeros(max(n, 1), np.uint32)
        sig, mus = [], []
        for k, t in enumerate(sets):
            if t is None:
                continue
            sg, mu = np.asarray(t[0], np.float64), np.asarray(t[1], np.float64)
            n_orient[k], stride[k] = sg.shape[0], mu.shape[1]
            sig.append(sg.ravel())
            mus.append(mu.ravel())
        sg = np.ascontiguousarray(np.concatenate(sig) if sig else np.zeros(1))
        mu = np.ascontiguousarray(np.concatenate(mus) if mus else np.zeros(1))
        u32 = C.POINTER(C.c_uint32)
        d = C.POINTER(C.c_double)
        _check(self.lib.icarion_gpu_set_ehss_offline(
            self.h, n_orient.ctypes.data_as(u32), stride.ctypes.data_as(u32),
            sg.ctypes.data_as(d), mu.ctypes.data_as(d), n), self.lib)

    def arrival_stats(self):
        """Count, min, max, mean and std of the resident ions' arrival times."""
        st = _abi.ArrivalStats()
        _check(self.lib.icarion_gpu_arrival_stats(self.h, C.byref(st)), self.lib)
        return st

    def arrival_histogram(self, lo, hi, n_bins):
        """Arrival-time counts in n_bins equal bins over [lo, hi] (numpy's rule)."""
        out = np.zeros(n_bins, dtype=np.uint64)
        _check(self.lib.icarion_gpu_arrival_histogram(self.h, lo, hi, n_bins,
                                                      out.ctypes.data_as(C.POINTER(C.c_uint64))),
               self.lib)
        return out

    def set_gas_radii(self, radii):
        """Contact radius per gas component of the mixture (0: the domain's)."""
        r = np.ascontiguousarray(radii, dtype=np.float64)
        _check(self.lib.icarion_gpu_set_gas_radii(self.h, r.ctypes.data_as(C.POINTER(C.c_double)),
                                                  len(r)), self.lib)

    def set_ehss_orientations(self, sets):
        """sets[s]: None (no orientation samples) or (quat_wxyz[n, 4], area_m2[n]);
        [] clears."""
        n = len(sets)
        n_orient = np.zeros(max(n, 1), np.uint32)
        qs, ars = [], []
        for k, t in enumerate(sets):
            if t is None:
                continue
            q, a = np.asarray(t[0], np.float64).reshape(-1, 4), np.asarray(t[1], np.float64).ravel()
            n_orient[k] = q.shape[0]
            qs.append(q.ravel())
            ars.append(a)
        q = np.ascontiguousarray(np.concatenate(qs) if qs else np.zeros(4))
        a = np.ascontiguousarray(np.concatenate(ars) if ars else np.zeros(1))
        d = C.POINTER(C.c_double)
        _check(self.lib.icarion_gpu_set_ehss_orientations(
            self.h, n_orient.ctypes.data_as(C.POINTER(C.c_uint32)), q.ctypes.data_as(d),
            a.ctypes.data_as(d), n), self.lib)

    def upload_field_terms(self, axes, terms):
        """axes: (xs, ys, zs) node coordinates (non-decreasing, possibly
        non-uniform); terms: [(E[nz, ny, nx, 3], scale_kind, constant_scale,
        frequency_Hz, phase_rad), ...] (E float32 or float64)."""
        xs, ys, zs = (np.ascontiguousarray(a, dtype=np.float64) for a in axes)
        arr = (_abi.FieldTermDesc * len(terms))()
        keep = []
        for i, (E, kind, const, freq, phase) in enumerate(terms):
            E = np.ascontiguousarray(E)
            assert E.shape[:3] == (len(zs), len(ys), len(xs))
            dtype = 0 if E.dtype == np.float32 else 1
            if dtype == 1:
                E = E.astype(np.float64, copy=False)
            keep.append(E)
            arr[i] = _abi.FieldTermDesc(E.ctypes.data_as(C.c_void_p), dtype, kind, const, freq,
                                        phase)
        d = C.POINTER(C.c_double)
        rc = self.lib.icarion_gpu_upload_field_terms(
            self.h, len(xs), len(ys), len(zs), xs.ctypes.data_as(d), ys.ctypes.data_as(d),
            zs.ctypes.data_as(d), arr, len(terms))
        _check(rc, self.lib)

    def upload_state(self, ens):
        s = ens.soa()
        _check(self.lib.icarion_gpu_upload_state(self.h, C.byref(s)), self.lib)

    def step(self, t0, dt, n_steps, step_offset=0):
        st = _abi.StepStats()
        _check(self.lib.icarion_gpu_step(self.h, t0, dt, n_steps, step_offset, C.byref(st)),
               self.lib)
        return st

    def step_noalloc(self, t0, dt, n_steps, step_offset=0):
        """icarion_gpu_step through the allocation-free entry point (the step
        table and counters in the library's fixed per-device scratch)."""
        st = _abi.StepStats()
        _check(self.lib.icarion_gpu_step_noalloc(self.h, t0, dt, n_steps, step_offset,
                                                 C.byref(st)), self.lib)
        return st

    def step_snapshots(self, t0, dt, n_steps, write_interval, step_offset=0, write_first=True,
                       next_write=None, on_snapshot=None, stop_after=None, copy=True):
        """icarion_gpu_step with trajectory snapshots.

        Snapshots follow SimulationEngine::run's sampling (SimulationEngine.cpp:275-278,
        OutputManager.cpp:230, 246-253); next_write defaults to write_interval, the value
        OutputManager::initialize sets (OutputManager.cpp:161).  Each snapshot is passed to
        on_snapshot(dict) when given, else collected; stop_after=q asks the library to stop
        after snapshot q.  copy=False hands on_snapshot views of the library's pinned
        buffers, valid during the call only.  Returns (snapshots, next_write, stats)."""
        nxt = C.c_double(write_interval if next_write is None else next_write)
        out = []

        def cb(_user, index, step, t, n, pos, vel, spec, dom, flags):
            if n:
                snap = {
                    "index": index, "step": step, "t": t,
                    "positions": np.ctypeslib.as_array(pos, shape=(n, 3)),
                    "velocities": np.ctypeslib.as_array(vel, shape=(n, 3)),
                    "species": np.ctypeslib.as_array(spec, shape=(n,)),
                    "domain": np.ctypeslib.as_array(dom, shape=(n,)),
                    "flags": np.ctypeslib.as_array(flags, shape=(n,)),
                }
                if copy or on_snapshot is None:
                    for k in ("positions", "velocities", "species", "domain", "flags"):
                        snap[k] = snap[k].copy()
            else:
                snap = {"index": index, "step": step, "t": t,
                        "positions": np.zeros((0, 3)), "velocities": np.zeros((0, 3)),
                        "species": np.zeros(0, np.uint32), "domain": np.zeros(0, np.int32),
                        "flags": np.zeros(0, np.uint8)}
            if on_snapshot is not None:
                on_snapshot(snap)
            else:
                out.append(snap)
            return 1 if (stop_after is not None and index >= stop_after) else 0

        fn = _abi.SnapshotFn(cb)
        st = _abi.StepStats()
        _check(self.lib.icarion_gpu_step_snapshots(self.h, t0, dt, n_steps, step_offset,
                                                   write_interval, 1 if write_first else 0,
                                                   C.byref(nxt), fn, None, C.byref(st)),
               self.lib)
        return out, nxt.value, st

    def download_state(self, ens):
        s = ens.soa()
        _check(self.lib.icarion_gpu_download_state(self.h, C.byref(s)), self.lib)
        return ens

    def synchronize(self):
        _check(self.lib.icarion_gpu_synchronize(self.h), self.lib)

    def step_batch(self, ens, t, dt, step_index):
        s = ens.soa()
        st = _abi.StepStats()
        _check(self.lib.icarion_gpu_step_batch(self.h, C.byref(s), t, dt, step_index,
                                               C.byref(st)), self.lib)
        return st


def device_count():
    return load().icarion_gpu_device_count()
