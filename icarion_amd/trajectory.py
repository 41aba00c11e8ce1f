"""Trajectory snapshots in the layout of the reference's /trajectory group.

OutputManager::log_step buffers one snapshot per sampled step
(OutputManager.cpp:169-231), and HDF5Writer::append_trajectory_batch_flat appends
the buffered snapshots to /trajectory (hdf5Writer.cpp:533-660):

    time                [T]        float64
    positions           [T, N, 3]  float64
    velocities          [T, N, 3]  float64
    domain_indices      [T, N]     int32
    species_id_indices  [T, N]     uint32

TrajectoryBuffer collects the snapshots that gpu.Handle.step_snapshots hands over
and can write the same datasets into an .npz file, with keys "trajectory/<name>".
H5TrajectoryFile writes the reference's HDF5 file itself (the /ions and
/trajectory groups, libicarion_h5.so, include/icarion_h5.h), fed by the same
snapshots.  domain_indices are the snapshot's domain indices, or 0 when it
carries none.

time_per_ion is not produced by the snapshot path: for the fixed-step
integrators the device keeps no per-ion time, and the reference writes that
dataset only when it has one (H5TrajectoryFile.append takes it when a caller
has it).
"""
import ctypes as C
import os

import numpy as np

from . import _abi

H5_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libicarion_h5.so")
_h5 = None


def load_h5():
    """libicarion_h5.so (built by icarion_amd.build next to the GPU library when
    an HDF5 C library is present); raises if absent."""
    global _h5
    if _h5 is not None:
        return _h5
    if not os.path.exists(H5_LIB_PATH):
        raise FileNotFoundError(f"{H5_LIB_PATH} not found: build it with `python -m icarion_amd.build` "
                                "on a machine with the HDF5 C library (HDF5_DIR)")
    lib = C.CDLL(H5_LIB_PATH)
    P, d = C.POINTER, C.POINTER(C.c_double)
    lib.icarion_h5_create.argtypes = [C.c_char_p, P(_abi.IonSoA), P(C.c_char_p), C.c_uint32, C.c_int,
                                      C.c_uint32, P(C.c_void_p)]
    lib.icarion_h5_append.argtypes = [C.c_void_p, C.c_double, d, d, P(C.c_int32), P(C.c_uint32), d]
    lib.icarion_h5_flush.argtypes = [C.c_void_p]
    lib.icarion_h5_update_death_times.argtypes = [C.c_void_p, d, C.c_size_t]
    lib.icarion_h5_close.argtypes = [C.c_void_p]
    lib.icarion_h5_rows.argtypes = [C.c_void_p]
    lib.icarion_h5_rows.restype = C.c_uint64
    lib.icarion_h5_last_error.argtypes = [C.c_char_p, C.c_size_t]
    lib.icarion_h5_last_error.restype = C.c_size_t
    for name in ("icarion_h5_create", "icarion_h5_append", "icarion_h5_flush",
                 "icarion_h5_update_death_times", "icarion_h5_close"):
        getattr(lib, name).restype = C.c_int
    _h5 = lib
    return lib


def _h5_check(rc, lib):
    if rc != 0:
        buf = C.create_string_buffer(4096)
        lib.icarion_h5_last_error(buf, 4096)
        raise RuntimeError(f"icarion_h5 error {rc}: {buf.value.decode(errors='replace')}")


class H5TrajectoryFile:
    """The reference's trajectory file: /ions from the initial ensemble at
    creation, then one /trajectory row per snapshot (a step_snapshots
    on_snapshot consumer), appended in batches of `batch`; death_times() at the
    end rewrites /ions/death_time_s (HDF5Writer::update_death_times)."""

    def __init__(self, path, ensemble, species_pool=None, deflate=2, batch=100):
        self.lib = load_h5()
        self.n = ensemble.n
        soa = ensemble.soa()
        pool = None
        if species_pool is not None:
            self._pool = [s.encode() for s in species_pool]
            pool = (C.c_char_p * len(self._pool))(*self._pool)
        h = C.c_void_p()
        _h5_check(self.lib.icarion_h5_create(os.fsencode(path), C.byref(soa), pool,
                                             len(species_pool) if species_pool else 0, deflate, batch,
                                             C.byref(h)), self.lib)
        self.h = h

    def __call__(self, snap):
        self.append(snap["t"], snap["positions"], snap["velocities"], snap["species"],
                    snap.get("domain"))

    def append(self, t, positions, velocities, species, domain=None, time_per_ion=None):
        d = C.POINTER(C.c_double)
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.n, 3)
        vel = np.ascontiguousarray(velocities, dtype=np.float64).reshape(self.n, 3)
        spec = np.ascontiguousarray(species, dtype=np.uint32).reshape(self.n)
        dom = None if domain is None else np.ascontiguousarray(domain, dtype=np.int32).reshape(self.n)
        tpi = None if time_per_ion is None else np.ascontiguousarray(time_per_ion, dtype=np.float64)
        _h5_check(self.lib.icarion_h5_append(
            self.h, float(t), pos.ctypes.data_as(d), vel.ctypes.data_as(d),
            dom.ctypes.data_as(C.POINTER(C.c_int32)) if dom is not None else C.POINTER(C.c_int32)(),
            spec.ctypes.data_as(C.POINTER(C.c_uint32)),
            tpi.ctypes.data_as(d) if tpi is not None else d()), self.lib)

    def rows(self):
        return int(self.lib.icarion_h5_rows(self.h))

    def flush(self):
        _h5_check(self.lib.icarion_h5_flush(self.h), self.lib)

    def death_times(self, death_time_s):
        a = np.ascontiguousarray(death_time_s, dtype=np.float64)
        _h5_check(self.lib.icarion_h5_update_death_times(self.h, a.ctypes.data_as(C.POINTER(C.c_double)),
                                                         a.size), self.lib)

    def close(self):
        if self.h:
            h, self.h = self.h, None
            _h5_check(self.lib.icarion_h5_close(h), self.lib)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class TrajectoryBuffer:
    """Appends snapshots and flushes them in batches, as OutputManager does.

    ``buffer_max`` snapshots are held before a flush (OutputManager.cpp:184-188,
    255-260).  ``sink(batch)`` receives each flushed batch as a dict of the
    datasets above; the default sink concatenates them in memory.  The buffer
    keeps copies of the snapshot arrays, so it may be fed the transient views
    of ``step_snapshots(copy=False)``.
    """

    def __init__(self, buffer_max=100, sink=None):
        if buffer_max <= 0:
            raise ValueError("buffer_max must be positive")
        self.buffer_max = buffer_max
        self.sink = sink
        self.n_ions = None
        self._buf = []
        self._flushed = []
        self.last_time = None

    def __call__(self, snap):
        self.add(snap)

    def add(self, snap):
        n = snap["positions"].shape[0]
        if self.n_ions is None:
            self.n_ions = n
        elif n != self.n_ions:
            raise RuntimeError("TrajectoryBuffer: ion count changed during the run")
        if len(self._buf) >= self.buffer_max:
            self.flush()
        self._buf.append({k: np.array(v) if isinstance(v, np.ndarray) else v for k, v in snap.items()})
        self.last_time = snap["t"]

    def _batch(self, snaps):
        n = self.n_ions or 0
        T = len(snaps)
        if T == 0:
            return {"time": np.zeros(0), "positions": np.zeros((0, n, 3)),
                    "velocities": np.zeros((0, n, 3)), "domain_indices": np.zeros((0, n), np.int32),
                    "species_id_indices": np.zeros((0, n), np.uint32)}
        return {
            "time": np.array([s["t"] for s in snaps], dtype=np.float64),
            "positions": np.stack([s["positions"] for s in snaps]).reshape(T, n, 3),
            "velocities": np.stack([s["velocities"] for s in snaps]).reshape(T, n, 3),
            "domain_indices": np.stack([s["domain"] if "domain" in s else np.zeros(n, np.int32)
                                        for s in snaps]).astype(np.int32).reshape(T, n),
            "species_id_indices": np.stack([s["species"] for s in snaps]).astype(np.uint32).reshape(T, n),
        }

    def flush(self):
        if not self._buf:
            return
        batch = self._batch(self._buf)
        self._buf = []
        if self.sink is not None:
            self.sink(batch)
        else:
            self._flushed.append(batch)

    def datasets(self):
        """Every snapshot so far, flushed or buffered, concatenated along T."""
        self.flush()
        if not self._flushed:
            return self._batch([])
        return {k: np.concatenate([b[k] for b in self._flushed]) for k in self._flushed[0]}

    def save_npz(self, path):
        d = self.datasets()
        np.savez(path, **{"trajectory/" + k: v for k, v in d.items()})
