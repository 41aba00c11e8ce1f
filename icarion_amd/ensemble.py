"""Host SoA ion ensemble (the hot/cold blocks of the reference IonEnsemble,
src/core/types/IonEnsemble.h:267-318) backed by numpy arrays, with a view as the
C-ABI struct icarion_ion_soa."""
import ctypes as C

import numpy as np

from . import _abi


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Ensemble:
    FIELDS_F64 = ("x", "y", "z", "vx", "vy", "vz", "mass_kg", "charge_C", "death_time_s",
                  "arrival_time_s", "time_s", "dt_s")

    def __init__(self, n):
        self.n = int(n)
        for f in self.FIELDS_F64:
            setattr(self, f, np.zeros(self.n, np.float64))
        self.species = np.zeros(self.n, np.uint32)
        self.active = np.ones(self.n, np.uint8)
        self.born = np.ones(self.n, np.uint8)
        self.birth_time_s = None
        self.n_collisions = np.zeros(self.n, np.uint32)
        self.death_time_s[:] = -1.0
        self.arrival_time_s[:] = np.nan
        self.time_s[:] = np.nan   # ion time t_i; NaN = the t0 of the first step

    def copy(self):
        e = Ensemble(self.n)
        for f in self.FIELDS_F64 + ("species", "active", "born", "n_collisions"):
            setattr(e, f, getattr(self, f).copy())
        e.birth_time_s = None if self.birth_time_s is None else self.birth_time_s.copy()
        return e

    def subset(self, idx):
        idx = np.asarray(idx)
        e = Ensemble(len(idx))
        for f in self.FIELDS_F64 + ("species", "active", "born", "n_collisions"):
            setattr(e, f, np.ascontiguousarray(getattr(self, f)[idx]))
        if self.birth_time_s is not None:
            e.birth_time_s = np.ascontiguousarray(self.birth_time_s[idx])
        return e

    def soa(self):
        """icarion_ion_soa pointing at this ensemble's arrays (keep self alive)."""
        d = C.c_double
        s = _abi.IonSoA()
        s.n = self.n
        for f in self.FIELDS_F64:
            setattr(s, f, _ptr(getattr(self, f), d))
        s.species = _ptr(self.species, C.c_uint32)
        s.active = _ptr(self.active, C.c_uint8)
        s.born = _ptr(self.born, C.c_uint8)
        s.birth_time_s = (_ptr(self.birth_time_s, d) if self.birth_time_s is not None
                          else C.POINTER(d)())
        s.n_collisions = _ptr(self.n_collisions, C.c_uint32)
        return s

    @property
    def pos(self):
        return np.stack([self.x, self.y, self.z], axis=1)

    @property
    def vel(self):
        return np.stack([self.vx, self.vy, self.vz], axis=1)

    @property
    def alive(self):
        return (self.active != 0) & (self.born != 0)
