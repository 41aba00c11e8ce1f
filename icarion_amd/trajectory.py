"""Trajectory snapshots in the layout of the reference's /trajectory group.

OutputManager::log_step buffers one snapshot per sampled step
(OutputManager.cpp:169-231), and HDF5Writer::append_trajectory_batch_flat appends
the buffered snapshots to /trajectory (hdf5Writer.cpp:533-660):

    time                [T]        float64
    positions           [T, N, 3]  float64
    velocities          [T, N, 3]  float64
    domain_indices      [T, N]     int32
    species_id_indices  [T, N]     uint32

TrajectoryBuffer collects the snapshots that gpu.Handle.step_snapshots hands over.
It writes the same datasets into an .npz file, with keys "trajectory/<name>".
This image has no HDF5 library, so HDF5 output itself is not provided.
domain_indices are the snapshot's domain indices, or 0 when it carries none.

time_per_ion is not produced. For the fixed-step integrators the device keeps no
per-ion time, and the reference writes that dataset only when it has one.
"""

import numpy as np


class TrajectoryBuffer:
    """Appends snapshots and flushes them in batches, as OutputManager does.

    ``buffer_max`` snapshots are held before a flush (OutputManager.cpp:184-188,
    255-260).  ``sink(batch)`` receives each flushed batch as a dict of the
    datasets above; the default sink concatenates them in memory.
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
        self._buf.append(snap)
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
