"""ctypes binding of libmccastro.so (declarations: include/mc_castro.h)."""
import ctypes as C
import os
import subprocess

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
DATA = os.path.join(PKG, "data")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


def lib_path():
    return os.path.join(LIBDIR, "libmccastro.so")


def table_path(kind="bin"):
    return os.path.join(DATA, "helm_table." + ("bin" if kind == "bin" else "dat"))


def build(jobs=8):
    """Compile the HIP library for gfx950 and generate the EOS table."""
    subprocess.run(["make", "-s", "-j%d" % jobs, "-C", CSRC], check=True)


class BatchBox(C.Structure):
    """mc_batch_box_t: a valid box and the device FABs (pointer, lo, hi) of
    S_old and S_new that hold it."""
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3),
                ("s_old", C.c_void_p), ("old_lo", C.c_int * 3), ("old_hi", C.c_int * 3),
                ("s_new", C.c_void_p), ("new_lo", C.c_int * 3), ("new_hi", C.c_int * 3)]


# (name, restype, argtypes) for every exported symbol of include/mc_castro.h
_I = C.c_int
_D = C.c_double
FAB = [_dp, _ip, _ip]
SIGS = [
    ("network_init", None, []),
    ("network_finalize", None, []),
    ("eos_init", None, []),
    ("eos_finalize", None, []),
    ("ctoprim", None, [_ip, _ip] + FAB * 3),
    ("enforce_minimum_density", None, [_ip, _ip] + FAB),
    ("normalize_species", None, [_ip, _ip] + FAB),
    ("reset_internal_e", None, [_ip, _ip] + FAB),
    ("compute_temp", None, [_ip, _ip] + FAB),
    ("estdt", None, [_ip, _ip] + FAB + [_dp, _dp]),
    ("store_flux", None, [_ip, _ip] + FAB * 3 + [_D]),
    ("fill_hydro_source", None, [_ip, _ip] + FAB * 13 + [_dp, _D]),
    ("normalize_species_fluxes", None, [_ip, _ip] + FAB),
    ("apply_av", None, [_ip, _ip, _I, _dp] + FAB * 3),
    ("trans1", None, [_ip, _ip, _I, _I] + FAB * 7 + [_D]),
    ("trans2", None, [_ip, _ip, _I, _I, _I] + FAB * 9 + [_D, _D, _D]),
    ("compute_flux", None, [_ip, _ip] + FAB * 6 + [_I]),
    ("divu", None, [_ip, _ip] + FAB + [_dp] + FAB),
    ("trace_ppm", None, [_ip, _ip, _ip, _ip, _I] + FAB * 4 + [_ip, _ip, _dp, _D]),
    ("initdata", None, [_ip, _ip] + FAB + [_dp, _dp, _dp]),
    ("denerror", None, [_ip, _ip, C.c_void_p, _ip, _ip] + FAB + [C.c_int8, C.c_int8]),
    ("calculate_blast_radius", None, [_ip, _ip] + FAB + [_dp, _dp, _dp, _dp, _dp, _D]),
    ("mc_calculate_blast_radius_noalloc", _I, [_ip, _ip] + FAB + [_dp, _dp, _dp, _dp, _dp, _D]),
    ("mc_set_stream", None, [_vp]),
    ("mc_get_stream", _vp, []),
    ("mc_last_error", C.c_char_p, []),
    ("mc_eos_init_from", _I, [C.c_char_p]),
    ("mc_eos_init_arrays", _I, [_dp]),
    ("mc_eos_init_generated", _I, []),
    ("mc_eos_eval_batch", _I, [_I, _I, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("mc_eos_newton_failures", _I, [C.POINTER(C.c_ulonglong), _I]),
    ("mc_set_problem", None, [_D, _D, _D, _D, _I]),
    ("mc_level_create", _vp, [_ip, _dp, _dp, _I, _I, _ip]),
    ("mc_level_destroy", None, [_vp]),
    ("mc_level_set_mode", _I, [_vp, _I]),
    ("mc_level_nbox", _I, [_vp]),
    ("mc_level_state_ptr", _dp, [_vp, _I]),
    ("mc_level_set_state", _I, [_vp, _I, _dp]),
    ("mc_level_get_state", _I, [_vp, _I, _dp]),
    ("mc_level_init_data", _I, [_vp, _D, _D, _D, _D, _I]),
    ("mc_level_est_time_step", _D, [_vp]),
    ("mc_level_initial_dt", _D, [_vp]),
    ("mc_level_advance", _I, [_vp, _D]),
    ("mc_level_post_timestep", _I, [_vp]),
    ("mc_level_coarse_step", _D, [_vp, _D]),
    ("mc_level_steps", _I, [_vp]),
    ("mc_level_time", _D, [_vp]),
    ("mc_level_dt", _D, [_vp]),
    ("mc_level_sync", _I, [_vp]),
    ("mc_level_run", _I, [_vp, _I, _D, _I]),
    ("mc_level_set_fused_slab", _I, [_vp, _I]),
    ("mc_level_fused_slab", _I, [_vp]),
    ("mc_level_set_tiling", _I, [_vp, _I, _I, _I, _I]),
    ("mc_level_get_tiling", _I, [_vp, _I, _ip]),
    ("mc_level_set_brick_order", _I, [_vp, _I, _I]),
    ("mc_level_brick_order", _I, [_vp, _I]),
    ("mc_level_run_stage", _I, [_vp, _I]),
    ("mc_level_autotune", _I, [_vp, _I, _ip, _dp]),
    ("mc_level_set_stage_variant", _I, [_vp, _I]),
    ("mc_level_stage_variant", _I, [_vp]),
    ("mc_level_run_timed", _I, [_vp, _I, _D, _dp, _ip]),
    ("mc_level_blast_radius", _I, [_vp, _dp, _dp, _dp]),
    ("mc_flux_register_create", _vp, [_I, _ip, _ip, _ip, _I]),
    ("mc_flux_register_destroy", None, [_vp]),
    ("mc_flux_register_size", C.c_long, [_vp]),
    ("mc_flux_register_get", _I, [_vp, _dp]),
    ("mc_flux_register_set_val", _I, [_vp, _D]),
    ("mc_flux_register_crse_init", _I, [_vp, _dp, _ip, _ip, _I, _D]),
    ("mc_flux_register_fine_add", _I, [_vp, _I, _dp, _ip, _ip, _I, _D]),
    ("mc_flux_register_clear_internal_borders", _I, [_vp]),
    ("mc_flux_register_reflux", _I, [_vp, _dp, _ip, _ip, _ip, _ip, _dp, _ip, _ip, _D]),
    ("mc_average_down", _I, [_dp, _ip, _ip, _dp, _ip, _ip, _ip, _ip, _ip, _I]),
    ("mc_level_set_flux_registers", _I, [_vp, _vp, _vp]),
    ("mc_level_reflux", _I, [_vp, _vp]),
    ("mc_level_average_down", _I, [_vp, _vp, _ip, _I, _ip]),
    ("mc_advance_batch", _I, [_ip, _dp, _dp, _I, _I, C.POINTER(BatchBox), _D, _dp]),
    ("mc_advance_batch_release", None, []),
    ("mc_ensemble_create", _vp, [_I, _ip, _dp, _dp, _ip]),
    ("mc_ensemble_create_slabs", _vp, [_I, _ip, _dp, _dp, _ip, _I]),
    ("mc_ensemble_slab", _I, [_vp]),
    ("mc_fused_bytes", C.c_long, [_I, _ip, _I]),
    ("mc_ensemble_destroy", None, [_vp]),
    ("mc_ensemble_ninst", _I, [_vp]),
    ("mc_ensemble_state_ptr", _dp, [_vp, _I]),
    ("mc_ensemble_init_data", _I, [_vp, _dp, _dp, _dp, _dp, _I]),
    ("mc_ensemble_set_state", _I, [_vp, _I, _dp]),
    ("mc_ensemble_get_state", _I, [_vp, _I, _dp]),
    ("mc_ensemble_run", _I, [_vp, _I, _D, _I]),
    ("mc_ensemble_clock", _I, [_vp, _I, _dp, _dp, _ip]),
    ("mc_ensemble_autotune", _I, [_vp, _I, _ip, _dp]),
    ("mc_ensemble_set_stage_variant", _I, [_vp, _I]),
    ("mc_ensemble_stage_variant", _I, [_vp]),
    ("mc_adjoint_create", _vp, [_ip, _dp, _dp, _I, _I]),
    ("mc_adjoint_destroy", None, [_vp]),
    ("mc_adjoint_forward", _I, [_vp, _vp, _vp, _I, _dp, _I]),
    ("mc_adjoint_backward", _I, [_vp, _vp, _vp]),
    ("mc_adjoint_steps", _I, [_vp]),
    ("mc_adjoint_keep_intermediates", _I, [_vp, _I]),
    ("mc_adjoint_keeps_intermediates", _I, [_vp]),
]

_lib = None


def lib():
    """Load libmccastro.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        p = lib_path()
        if not os.path.exists(p):
            raise RuntimeError("libmccastro.so is not built (run make -C %s)" % CSRC)
        L = C.CDLL(p)
        for name, res, args in SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        err = lib().mc_last_error()
        raise RuntimeError("%s failed (%d): %s" % (what, rc, err.decode() if err else ""))


_table_loaded = False


def newton_failures(reset=False):
    """EOS Newton solves that ended unconverged since the last reset."""
    v = C.c_ulonglong(0)
    check(lib().mc_eos_newton_failures(C.byref(v), int(bool(reset))), "mc_eos_newton_failures")
    return int(v.value)


def load_table(path=None):
    """Upload the EOS table to the device (generated binary cache by default)."""
    global _table_loaded
    if path is None:
        path = table_path("bin")
        if not os.path.exists(path):
            check(lib().mc_eos_init_generated(), "mc_eos_init_generated")
            _table_loaded = True
            return
    check(lib().mc_eos_init_from(path.encode()), "mc_eos_init_from")
    _table_loaded = True
