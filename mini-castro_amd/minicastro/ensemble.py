"""Ensembles of independent Sedov instances advanced together (mc_ensemble_*)."""
import ctypes as C

import numpy as np

from ._lib import lib, check, _dp, _ip

NVAR = 20


class Ensemble:
    """instances: list of dicts with n_cell, prob_lo, prob_hi, bc and the
    problem constants p_ambient, dens_ambient, exp_energy, r_init, nsub."""

    def __init__(self, instances, slab=-1):
        """slab: z-slab planes of the fused engine's face arrays (-1: chosen
        by free memory, 0: whole instances)."""
        self.L = lib()
        self.inst = list(instances)
        m = len(self.inst)
        nc = np.ascontiguousarray([x for d in self.inst for x in d["n_cell"]], dtype=np.int32)
        lo = np.ascontiguousarray([x for d in self.inst for x in d["prob_lo"]], dtype=np.float64)
        hi = np.ascontiguousarray([x for d in self.inst for x in d["prob_hi"]], dtype=np.float64)
        bc = np.ascontiguousarray([d["bc"] for d in self.inst], dtype=np.int32)
        self.h = self.L.mc_ensemble_create_slabs(m, nc.ctypes.data_as(_ip), lo.ctypes.data_as(_dp),
                                                 hi.ctypes.data_as(_dp), bc.ctypes.data_as(_ip), int(slab))
        if not self.h:
            raise RuntimeError("mc_ensemble_create: %s" % self.L.mc_last_error().decode())

    @property
    def slab(self):
        """z-slab planes in use (0: whole instances)."""
        return self.L.mc_ensemble_slab(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.L.mc_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_data(self):
        arr = lambda k: np.ascontiguousarray([d[k] for d in self.inst], dtype=np.float64)
        pa, da, ea, ra = arr("p_ambient"), arr("dens_ambient"), arr("exp_energy"), arr("r_init")
        check(self.L.mc_ensemble_init_data(self.h, pa.ctypes.data_as(_dp), da.ctypes.data_as(_dp),
                                           ea.ctypes.data_as(_dp), ra.ctypes.data_as(_dp),
                                           int(self.inst[0].get("nsub", 10))), "mc_ensemble_init_data")

    def set_stage_variant(self, flags):
        """Kernel variant flags of the fused step (minicastro.Level.SPLIT_TRACE,
        FACE_TRANS1, CELL_FINAL); results are identical."""
        check(self.L.mc_ensemble_set_stage_variant(self.h, int(flags)), "mc_ensemble_set_stage_variant")

    @property
    def stage_variant(self):
        return self.L.mc_ensemble_stage_variant(self.h)

    def run(self, nsteps, stop_time=-1.0, use_graph=True):
        check(self.L.mc_ensemble_run(self.h, int(nsteps), stop_time, 1 if use_graph else 0), "mc_ensemble_run")

    def state(self, m):
        n = tuple(self.inst[m]["n_cell"])
        a = np.zeros(n + (NVAR,), order="F")
        check(self.L.mc_ensemble_get_state(self.h, m, a.ctypes.data_as(_dp)), "mc_ensemble_get_state")
        return a

    def clock(self, m):
        dt, t = C.c_double(), C.c_double()
        s = C.c_int()
        check(self.L.mc_ensemble_clock(self.h, m, C.byref(dt), C.byref(t), C.byref(s)), "mc_ensemble_clock")
        return dt.value, t.value, s.value

    def autotune(self, reps=3):
        """Pick the fastest brick of every stage (state unchanged).
        Returns {stage: ((bx, by, bz), ms_per_launch)}."""
        from .level import Level, _autotune
        return _autotune(self.L.mc_ensemble_autotune, self.h, reps, Level.STAGES, "mc_ensemble_autotune")

    def zones(self):
        return sum(int(np.prod(d["n_cell"])) for d in self.inst)
