"""ORACLE — test infrastructure only.

ctypes binding of the CPU restatement of mini-Castro's CTU path
(``oracle/build/liboracle.so``, built by ``oracle/Makefile``).  Only the tests,
the smoke check and the benchmark's CPU-baseline leg import this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORC_LIB names another build of the same sources (the pin tests are run
# against deliberately altered builds to show that they catch the change)
LIB_PATH = os.environ.get("ORC_LIB") or os.path.join(HERE, "build", "liboracle.so")

NVAR, QVAR, NQAUX, NGDNV, NSPEC = 20, 21, 4, 6, 13
BC_PERIODIC, BC_OUTFLOW = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class EosT(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "rho T p e h s abar zbar dpdT dpdr dedT dedr dhdT dhdr dsdT dsdr dpde dpdr_e "
        "cv cp xne xnp eta pele ppos mu mu_e y_e gam1 cs dpdA dpdZ dedA dedZ").split()] + [
        ("iat", C.c_int), ("jat", C.c_int), ("niter", C.c_int), ("pad_", C.c_int)]


class ProbT(C.Structure):
    _fields_ = [("p_ambient", C.c_double), ("dens_ambient", C.c_double),
                ("exp_energy", C.c_double), ("r_init", C.c_double),
                ("nsub", C.c_int), ("pad_", C.c_int)]


def build():
    """Compile the oracle (make in oracle/)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def _bind(path):
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    if True:
        L.orc_level_create.restype = C.c_void_p
        L.orc_level_state.restype = _dp
        L.orc_level_source.restype = _dp
        L.orc_level_est_time_step.restype = C.c_double
        L.orc_level_initial_dt.restype = C.c_double
        L.orc_level_coarse_step.restype = C.c_double
        L.orc_level_time.restype = C.c_double
        L.orc_level_dt.restype = C.c_double
        for f in ("orc_level_state", "orc_level_source", "orc_level_est_time_step",
                  "orc_level_initial_dt", "orc_level_coarse_step", "orc_level_time",
                  "orc_level_dt", "orc_level_nbox", "orc_level_steps", "orc_level_destroy",
                  "orc_level_init_data", "orc_level_advance", "orc_level_post_timestep",
                  "orc_level_clean_state_new"):
            getattr(L, f).argtypes = [C.c_void_p] + ([C.c_int] if f in ("orc_level_state", "orc_level_source") else [])
        L.orc_level_coarse_step.argtypes = [C.c_void_p, C.c_double]
        L.orc_level_advance.argtypes = [C.c_void_p, C.c_double]
        L.orc_level_init_data.argtypes = [C.c_void_p, C.POINTER(ProbT)]
        L.orc_eos.argtypes = [C.c_int, C.POINTER(EosT)]
        L.orc_eos_init.argtypes = [C.c_char_p]
        L.orc_ppm_reconstruct.argtypes = [_dp, C.c_double, _dp, _dp]
        L.orc_ppm_int_profile.argtypes = [C.c_double] * 6 + [_dp, _dp]
        _bind_amr(L)
    return L


def _bind_amr(L):
    vp, I, D = C.c_void_p, C.c_int, C.c_double
    L.orc_flux_register_create.restype = vp
    L.orc_flux_register_create.argtypes = [I, _ip, _ip, _ip, I]
    L.orc_flux_register_destroy.argtypes = [vp]
    L.orc_flux_register_size.restype = C.c_long
    L.orc_flux_register_size.argtypes = [vp]
    L.orc_flux_register_data.restype = _dp
    L.orc_flux_register_data.argtypes = [vp]
    L.orc_flux_register_set_val.argtypes = [vp, D]
    L.orc_flux_register_crse_init.argtypes = [vp, _dp, _ip, _ip, I, D]
    L.orc_flux_register_fine_add.argtypes = [vp, I, _dp, _ip, _ip, I, D]
    L.orc_flux_register_clear_internal_borders.argtypes = [vp]
    L.orc_flux_register_reflux.argtypes = [vp, _dp, _ip, _ip, _ip, _ip, _dp, _ip, _ip, D]
    L.orc_average_down.argtypes = [_dp, _ip, _ip, _dp, _ip, _ip, _ip, _ip, _ip, I]
    L.orc_level_set_flux_registers.argtypes = [vp, vp, vp]
    L.orc_level_reflux.argtypes = [vp, vp]
    L.orc_level_average_down.argtypes = [vp, vp, _ip, I, _ip]


def register_faces(fine_boxes, ratio):
    """(fine box, dir, side, face lo, face hi) of every register face, in the
    register's storage order (per fine box, direction, side)."""
    out = []
    for b, fb in enumerate(np.asarray(fine_boxes).reshape(-1, 6)):
        clo = [int(fb[d]) // ratio[d] for d in range(3)]
        chi = [(int(fb[3 + d]) + 1) // ratio[d] - 1 for d in range(3)]
        for d in range(3):
            for side in (0, 1):
                lo, hi = list(clo), list(chi)
                lo[d] = hi[d] = chi[d] + 1 if side else clo[d]
                out.append((b, d, side, tuple(lo), tuple(hi)))
    return out


class FluxRegister:
    """The oracle's flux register (oracle/orc_amr.c)."""

    def __init__(self, fine_boxes, ratio, crse_n, bc, _library=None):
        self.L = _library or lib()
        fb = np.ascontiguousarray(np.asarray(fine_boxes, dtype=np.int32).reshape(-1, 6))
        self.fine_boxes, self.ratio = fb, tuple(int(r) for r in ratio)
        self.h = self.L.orc_flux_register_create(len(fb), fb.ctypes.data_as(_ip), iv(*ratio), iv(*crse_n), int(bc))
        self.faces = register_faces(fb, self.ratio)

    def __del__(self):
        try:
            self.L.orc_flux_register_destroy(self.h)
        except Exception:
            pass

    def data(self):
        n = self.L.orc_flux_register_size(self.h)
        return np.ctypeslib.as_array(self.L.orc_flux_register_data(self.h), shape=(n,))

    def face_arrays(self):
        """{(fine box, dir, side): (lo, hi, array of shape face + (NVAR,))}"""
        d, off, out = self.data(), 0, {}
        for b, dr, side, lo, hi in self.faces:
            shp = tuple(h - l + 1 for l, h in zip(lo, hi))
            n = int(np.prod(shp))
            out[(b, dr, side)] = (lo, hi, d[off:off + n * NVAR].reshape(shp + (NVAR,), order="F"))
            off += n * NVAR
        return out

    def set_val(self, v):
        self.L.orc_flux_register_set_val(self.h, float(v))

    def crse_init(self, fab, d, mult):
        self.L.orc_flux_register_crse_init(self.h, *fab.args(), int(d), float(mult))

    def fine_add(self, j, fab, d, mult):
        self.L.orc_flux_register_fine_add(self.h, int(j), *fab.args(), int(d), float(mult))

    def clear_internal_borders(self):
        self.L.orc_flux_register_clear_internal_borders(self.h)

    def reflux(self, state, lo, hi, vol, scale=1.0):
        self.L.orc_flux_register_reflux(self.h, *state.args(), iv(*lo), iv(*hi), *vol.args(), float(scale))


def average_down(fine, crse, lo, hi, ratio, ncomp=None):
    lib().orc_average_down(*fine.args(), *crse.args(), iv(*lo), iv(*hi), iv(*ratio),
                           int(ncomp if ncomp is not None else fine.nc))


def lib():
    global _lib
    if _lib is None:
        _lib = _bind(LIB_PATH)
    return _lib


def ptr(a):
    assert a.dtype == np.float64 and a.flags["F_CONTIGUOUS"] or a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iv(*v):
    return (C.c_int * 3)(*v)


def dv(*v):
    return (C.c_double * 3)(*v)


# ---------------------------------------------------------------- EOS
_eos_loaded = None


def eos_init(path):
    global _eos_loaded
    if _eos_loaded != path:
        rc = lib().orc_eos_init(path.encode())
        if rc != 0:
            raise RuntimeError("orc_eos_init(%s) failed: %d" % (path, rc))
        _eos_loaded = path


def eos_init_arrays(tab):
    """tab: (21, 201, 541) float64 (column, j, i)."""
    global _eos_loaded
    tab = np.ascontiguousarray(tab, dtype=np.float64)
    lib().orc_eos_init_arrays(ptr(tab))
    _eos_loaded = None


def eos_grid():
    t = np.zeros(201)
    d = np.zeros(541)
    lib().orc_eos_grid(ptr(t), ptr(d))
    return t, d


EOS_RT, EOS_RE, EOS_RP = 1, 2, 3


def eos(mode, rho, T, abar, zbar, e=0.0, p=0.0):
    st = EosT()
    st.rho, st.T, st.abar, st.zbar, st.e, st.p = rho, T, abar, zbar, e, p
    lib().orc_eos(mode, C.byref(st))
    return st


# ---------------------------------------------------------------- FABs
class Fab:
    """Fortran-ordered array on box [lo, hi] with nc components."""

    def __init__(self, lo, hi, nc, data=None):
        self.lo = tuple(int(x) for x in lo)
        self.hi = tuple(int(x) for x in hi)
        self.nc = nc
        shape = tuple(h - l + 1 for l, h in zip(self.lo, self.hi)) + (nc,)
        if data is None:
            self.a = np.zeros(shape, dtype=np.float64, order="F")
        else:
            self.a = np.asfortranarray(np.asarray(data, dtype=np.float64).reshape(shape, order="F"))

    def args(self):
        return (self.a.ctypes.data_as(_dp), iv(*self.lo), iv(*self.hi))


def call(name, *args):
    """Call an oracle kernel; Fab arguments expand to (ptr, lo, hi)."""
    fl = []
    for a in args:
        if isinstance(a, Fab):
            fl.extend(a.args())
        elif isinstance(a, float):
            fl.append(C.c_double(a))
        elif isinstance(a, (int, np.integer)):
            fl.append(C.c_int(int(a)))
        else:
            fl.append(a)
    getattr(lib(), "orc_" + name)(*fl)


# ---------------------------------------------------------------- level
class Level:
    def __init__(self, n_cell, prob_lo, prob_hi, bc, boxes, nthreads=1, _library=None):
        self.L = _library or lib()
        boxes = np.asarray(boxes, dtype=np.int32).reshape(-1, 6)
        self.boxes = boxes
        bx = np.ascontiguousarray(boxes.ravel())
        self.h = self.L.orc_level_create(iv(*n_cell), dv(*prob_lo), dv(*prob_hi), int(bc),
                                         len(boxes), bx.ctypes.data_as(_ip), int(nthreads))
        self.n_cell = tuple(n_cell)

    def __del__(self):
        try:
            self.L.orc_level_destroy(self.h)
        except Exception:
            pass

    def init_data(self, p_ambient, dens_ambient, exp_energy, r_init, nsub=10):
        pr = ProbT(p_ambient, dens_ambient, exp_energy, r_init, nsub, 0)
        self.L.orc_level_init_data(self.h, C.byref(pr))

    def box_state(self, b):
        lo, hi = self.boxes[b][:3], self.boxes[b][3:]
        n = int(np.prod(hi - lo + 1)) * NVAR
        p = self.L.orc_level_state(self.h, b)
        return np.ctypeslib.as_array(p, shape=(n,))

    def state(self):
        """Whole-domain state, shape (nx, ny, nz, NVAR) (Fortran order)."""
        out = np.zeros(self.n_cell + (NVAR,), order="F")
        for b in range(len(self.boxes)):
            lo, hi = self.boxes[b][:3], self.boxes[b][3:]
            shp = tuple(hi - lo + 1) + (NVAR,)
            out[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1, :] = \
                self.box_state(b).reshape(shp, order="F")
        return out

    def set_state(self, S):
        for b in range(len(self.boxes)):
            lo, hi = self.boxes[b][:3], self.boxes[b][3:]
            v = self.box_state(b)
            v[:] = np.asfortranarray(S[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1, :]).ravel(order="F")

    def est_time_step(self):
        return self.L.orc_level_est_time_step(self.h)

    def initial_dt(self):
        return self.L.orc_level_initial_dt(self.h)

    def advance(self, dt):
        self.L.orc_level_advance(self.h, dt)

    def post_timestep(self):
        self.L.orc_level_post_timestep(self.h)

    def coarse_step(self, stop_time=-1.0):
        return self.L.orc_level_coarse_step(self.h, stop_time)

    def set_flux_registers(self, crse_side=None, fine_side=None):
        self._regs = (crse_side, fine_side)   # kept alive with the level
        self.L.orc_level_set_flux_registers(self.h, crse_side.h if crse_side else None,
                                            fine_side.h if fine_side else None)

    def reflux(self, R):
        self.L.orc_level_reflux(self.h, R.h)

    def average_down_from(self, fine, ratio, fine_box_ids=None):
        ids = None if fine_box_ids is None else np.ascontiguousarray(fine_box_ids, dtype=np.int32)
        self.L.orc_level_average_down(fine.h, self.h, iv(*ratio), -1 if ids is None else len(ids),
                                      None if ids is None else ids.ctypes.data_as(_ip))

    @property
    def steps(self):
        return self.L.orc_level_steps(self.h)

    @property
    def time(self):
        return self.L.orc_level_time(self.h)

    @property
    def dt(self):
        return self.L.orc_level_dt(self.h)


# ------------------------------------------------------------ tangent-linear
# build/liboracle_tl.so: the same restatement with real = (value, derivative)
# dual numbers.  A state array holds NVAR x cells duals, i.e. pairs of doubles.
TL_PATH = os.path.join(HERE, "build", "liboracle_tl.so")


class DualT(C.Structure):
    _fields_ = [("v", C.c_double), ("d", C.c_double)]


_tl = None


def tl_lib():
    global _tl
    if _tl is None:
        if not os.path.exists(TL_PATH):
            build()
        L = C.CDLL(TL_PATH)
        L.orc_level_create.restype = C.c_void_p
        L.orc_level_create.argtypes = [_ip, _dp, _dp, C.c_int, C.c_int, _ip, C.c_int]
        L.orc_level_state.restype = _dp
        L.orc_level_state.argtypes = [C.c_void_p, C.c_int]
        L.orc_level_advance.argtypes = [C.c_void_p, DualT]
        L.orc_level_post_timestep.argtypes = [C.c_void_p]
        L.orc_level_destroy.argtypes = [C.c_void_p]
        L.orc_eos_init_arrays.argtypes = [_dp]
        _tl = L
    return _tl


def tl_eos_init_arrays(tab):
    t = np.ascontiguousarray(tab, dtype=np.float64)
    tl_lib().orc_eos_init_arrays(t.ctypes.data_as(_dp))


class TLLevel:
    """Tangent-linear of the oracle level: states carry one directional
    derivative.  advance(dt) + post_timestep() is one step with a given dt."""

    def __init__(self, n_cell, prob_lo, prob_hi, bc, boxes, nthreads=1):
        self.L = tl_lib()
        boxes = np.asarray(boxes, dtype=np.int32).reshape(-1, 6)
        self.boxes = boxes
        bx = np.ascontiguousarray(boxes.ravel())
        lo = np.ascontiguousarray(np.stack([np.asarray(prob_lo, float), np.zeros(3)], axis=1))
        hi = np.ascontiguousarray(np.stack([np.asarray(prob_hi, float), np.zeros(3)], axis=1))
        self.h = self.L.orc_level_create(iv(*n_cell), lo.ctypes.data_as(_dp), hi.ctypes.data_as(_dp), int(bc),
                                         len(boxes), bx.ctypes.data_as(_ip), int(nthreads))
        self.n_cell = tuple(n_cell)

    def __del__(self):
        try:
            self.L.orc_level_destroy(self.h)
        except Exception:
            pass

    def _box(self, b):
        lo, hi = self.boxes[b][:3], self.boxes[b][3:]
        n = int(np.prod(hi - lo + 1)) * NVAR
        return np.ctypeslib.as_array(self.L.orc_level_state(self.h, b), shape=(n, 2))

    def set_state(self, S, dS):
        for b in range(len(self.boxes)):
            lo, hi = self.boxes[b][:3], self.boxes[b][3:]
            v = self._box(b)
            sl = (slice(lo[0], hi[0] + 1), slice(lo[1], hi[1] + 1), slice(lo[2], hi[2] + 1))
            v[:, 0] = np.asfortranarray(S[sl]).ravel(order="F")
            v[:, 1] = np.asfortranarray(dS[sl]).ravel(order="F")

    def state(self):
        out = np.zeros(self.n_cell + (NVAR,), order="F")
        dout = np.zeros(self.n_cell + (NVAR,), order="F")
        for b in range(len(self.boxes)):
            lo, hi = self.boxes[b][:3], self.boxes[b][3:]
            shp = tuple(hi - lo + 1) + (NVAR,)
            v = self._box(b)
            sl = (slice(lo[0], hi[0] + 1), slice(lo[1], hi[1] + 1), slice(lo[2], hi[2] + 1))
            out[sl] = v[:, 0].reshape(shp, order="F")
            dout[sl] = v[:, 1].reshape(shp, order="F")
        return out, dout

    def step(self, dt):
        self.L.orc_level_advance(self.h, DualT(float(dt), 0.0))
        self.L.orc_level_post_timestep(self.h)


# ------------------------------------------------------------------ op count
# build/liboracle_cnt.so: the same restatement, single-threaded, with a scalar
# that counts its floating-point operations (values equal the double build).
CNT_PATH = os.path.join(HERE, "build", "liboracle_cnt.so")
OPCOUNT_FIELDS = ("add", "mul", "div", "sqrt", "log")
_cnt = None


def cnt_lib():
    global _cnt
    if _cnt is None:
        _cnt = _bind(CNT_PATH)
        _cnt.orc_opcount_read.argtypes = [C.POINTER(C.c_ulonglong)]
    return _cnt


class CntLevel(Level):
    """Oracle level of the op-count build; opcount() reads the operations
    executed since the last opcount_reset()."""

    def __init__(self, n_cell, prob_lo, prob_hi, bc, boxes):
        super().__init__(n_cell, prob_lo, prob_hi, bc, boxes, nthreads=1, _library=cnt_lib())

    def opcount_reset(self):
        self.L.orc_opcount_reset()

    def opcount(self):
        out = (C.c_ulonglong * 5)()
        self.L.orc_opcount_read(out)
        return dict(zip(OPCOUNT_FIELDS, (int(v) for v in out)))


def cnt_eos_init_arrays(tab):
    t = np.ascontiguousarray(tab, dtype=np.float64)
    cnt_lib().orc_eos_init_arrays(t.ctypes.data_as(_dp))


# ------------------------------------------------------------------ timing
# build/liboracle_fast.so: the same restatement built for speed (-O3, AVX2,
# FMA contraction allowed); its results may differ in the last bits, so only
# the CPU-baseline timing uses it.
FAST_PATH = os.path.join(HERE, "build", "liboracle_fast.so")
FAST_FLAGS = "-O3 -march=x86-64-v3 -ffp-contract=fast"
_fast = None


def fast_lib():
    global _fast
    if _fast is None:
        _fast = _bind(FAST_PATH)
    return _fast
