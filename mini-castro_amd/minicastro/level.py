"""One AMR level of boxes advanced on the GPU through the C ABI
(mc_level_*: the role of Castro::advance / coarseTimeStep)."""
import ctypes as C

import numpy as np

from ._lib import lib, check, _dp, _ip

MODE_PLAIN, MODE_FUSED = 0, 1
BC_PERIODIC, BC_OUTFLOW = 0, 1
NVAR = 20


class Level:
    def __init__(self, n_cell, boxes, prob_lo=(0.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 1.0), bc=BC_OUTFLOW,
                 mode=MODE_PLAIN):
        self.L = lib()
        self.n_cell = tuple(int(x) for x in n_cell)
        self.boxes = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        nc = (C.c_int * 3)(*self.n_cell)
        lo = (C.c_double * 3)(*prob_lo)
        hi = (C.c_double * 3)(*prob_hi)
        self.h = self.L.mc_level_create(nc, lo, hi, int(bc), len(self.boxes),
                                        self.boxes.ctypes.data_as(_ip))
        if not self.h:
            raise RuntimeError("mc_level_create failed: %s" % self.L.mc_last_error().decode())
        self.set_mode(mode)

    def close(self):
        if getattr(self, "h", None):
            self.L.mc_level_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mode(self, mode):
        check(self.L.mc_level_set_mode(self.h, int(mode)), "mc_level_set_mode")
        self.mode = mode

    def init_data(self, p_ambient, dens_ambient, exp_energy, r_init, nsub=10, **_):
        check(self.L.mc_level_init_data(self.h, p_ambient, dens_ambient, exp_energy, r_init, int(nsub)),
              "mc_level_init_data")

    def _shape(self, b):
        lo, hi = self.boxes[b][:3], self.boxes[b][3:]
        return tuple(int(x) for x in (hi - lo + 1)) + (NVAR,)

    def box_state(self, b):
        a = np.zeros(self._shape(b), order="F")
        check(self.L.mc_level_get_state(self.h, b, a.ctypes.data_as(_dp)), "mc_level_get_state")
        return a

    def state(self):
        out = np.zeros(self.n_cell + (NVAR,), order="F")
        for b in range(len(self.boxes)):
            lo, hi = self.boxes[b][:3], self.boxes[b][3:]
            out[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1, :] = self.box_state(b)
        return out

    def set_state(self, S):
        for b in range(len(self.boxes)):
            lo, hi = self.boxes[b][:3], self.boxes[b][3:]
            a = np.asfortranarray(S[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1, :])
            check(self.L.mc_level_set_state(self.h, b, a.ctypes.data_as(_dp)), "mc_level_set_state")

    def est_time_step(self):
        v = self.L.mc_level_est_time_step(self.h)
        if v < 0:
            raise RuntimeError("mc_level_est_time_step: %s" % self.L.mc_last_error().decode())
        return v

    def initial_dt(self):
        """Castro::computeInitialDt: init_shrink * estTimeStep."""
        v = self.L.mc_level_initial_dt(self.h)
        if v < 0:
            raise RuntimeError("mc_level_initial_dt: %s" % self.L.mc_last_error().decode())
        return v

    def advance(self, dt):
        check(self.L.mc_level_advance(self.h, dt), "mc_level_advance")

    def post_timestep(self):
        check(self.L.mc_level_post_timestep(self.h), "mc_level_post_timestep")

    def coarse_step(self, stop_time=-1.0):
        v = self.L.mc_level_coarse_step(self.h, stop_time)
        if v < 0:
            raise RuntimeError("mc_level_coarse_step: %s" % self.L.mc_last_error().decode())
        return v

    def run(self, nsteps, stop_time=-1.0, use_graph=True):
        """nsteps coarse steps; in fused mode dt stays on the device and the
        step is replayed from a captured hipGraph."""
        check(self.L.mc_level_run(self.h, int(nsteps), stop_time, 1 if use_graph else 0), "mc_level_run")

    # fused stages: "tflux" is "trace" and "flux1" in one launch, "trans1c"
    # and "finalc" are "trans1" and "final" by cells (which ones a step runs:
    # set_stage_variant); "dt" the time-step control kernels
    STAGES = ("est", "prim", "flat", "trace", "flux1", "trans1", "final", "update", "tflux", "trans1c", "finalc",
              "dt")
    SPLIT_TRACE, FACE_TRANS1, CELL_FINAL = 1, 2, 4

    def run_timed(self, nsteps, stop_time=-1.0):
        """Fused mode: nsteps eager steps, HIP events around each stage launch.
        Returns {stage: (total_ms, launches)}; a stage the step does not run
        has 0 launches."""
        n = len(self.STAGES)
        ms = (C.c_double * n)()
        nl = (C.c_int * n)()
        check(self.L.mc_level_run_timed(self.h, int(nsteps), stop_time, ms, nl), "mc_level_run_timed")
        return {self.STAGES[i]: (ms[i], nl[i]) for i in range(n)}

    def set_stage_variant(self, flags):
        """Fused mode: kernel variants, bit flags (0, the default: tflux,
        trans1c, final).  SPLIT_TRACE runs trace + flux1 instead of tflux,
        FACE_TRANS1 trans1 instead of trans1c, CELL_FINAL finalc instead of
        final; results are identical."""
        check(self.L.mc_level_set_stage_variant(self.h, int(flags)), "mc_level_set_stage_variant")

    @property
    def stage_variant(self):
        """The kernel variant flags in use (the autotuner sets them)."""
        v = self.L.mc_level_stage_variant(self.h)
        if v < 0:
            raise RuntimeError("mc_level_stage_variant: %s" % self.L.mc_last_error().decode())
        return v

    def set_fused_slab(self, planes):
        """z-slab planes of the fused engine's face arrays: -1 chosen by free
        memory (the default), 0 the whole level, > 0 that many cell planes."""
        check(self.L.mc_level_set_fused_slab(self.h, int(planes)), "mc_level_set_fused_slab")

    @property
    def fused_slab(self):
        v = self.L.mc_level_fused_slab(self.h)
        if v < 0:
            raise RuntimeError("mc_level_fused_slab: %s" % self.L.mc_last_error().decode())
        return v

    def set_tiling(self, stage, bx, by, bz):
        check(self.L.mc_level_set_tiling(self.h, int(stage), bx, by, bz), "mc_level_set_tiling")

    def get_tiling(self, stage):
        a = (C.c_int * 3)()
        check(self.L.mc_level_get_tiling(self.h, int(stage), a), "mc_level_get_tiling")
        return tuple(a)

    def set_brick_order(self, stage, zgroup):
        """0: raster brick order; 2..64: groups of that many brick planes in z."""
        check(self.L.mc_level_set_brick_order(self.h, int(stage), int(zgroup)), "mc_level_set_brick_order")

    def brick_order(self, stage):
        v = self.L.mc_level_brick_order(self.h, int(stage))
        if v < 0:
            raise RuntimeError("mc_level_brick_order: %s" % self.L.mc_last_error().decode())
        return v

    def run_stage(self, stage):
        check(self.L.mc_level_run_stage(self.h, int(stage)), "mc_level_run_stage")

    def autotune(self, reps=3):
        """Fused mode: pick the fastest brick of every stage and the faster
        kernel of each alternative (state unchanged; stage_variant reports
        the choice).  Returns {stage: ((bx, by, bz), ms_per_launch)}."""
        return _autotune(self.L.mc_level_autotune, self.h, reps, self.STAGES, "mc_level_autotune")

    def blast_radius(self):
        """(blast_mass, blast_radius_sum, max_density) of post_timestep's diagnostic."""
        m, r, x = C.c_double(), C.c_double(), C.c_double()
        check(self.L.mc_level_blast_radius(self.h, C.byref(m), C.byref(r), C.byref(x)), "mc_level_blast_radius")
        return m.value, r.value, x.value

    def sync(self):
        check(self.L.mc_level_sync(self.h), "mc_level_sync")

    @property
    def steps(self):
        return self.L.mc_level_steps(self.h)

    @property
    def time(self):
        return self.L.mc_level_time(self.h)

    @property
    def dt(self):
        return self.L.mc_level_dt(self.h)


def _autotune(fn, h, reps, names, what):
    """{stage: ((bx, by, bz), ms_per_launch)} of the stages the autotuner
    timed (the stages a step runs)."""
    n = len(names) - 1
    b = (C.c_int * (3 * n))()
    ms = (C.c_double * n)()
    check(fn(h, int(reps), b, ms), what)
    return {names[s]: ((b[3 * s], b[3 * s + 1], b[3 * s + 2]), ms[s]) for s in range(1, n) if ms[s] > 0}
