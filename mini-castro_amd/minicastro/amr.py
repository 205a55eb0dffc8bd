"""AMR coupling of a coarse and a fine level through the C ABI: the flux
register of a fine level (mc_flux_register_*: the reference's
FluxRegister CrseInit / FineAdd at Source/Castro_advance.cpp:82-88,
Castro::reflux at Source/Castro.cpp:451-486) and Castro::avgDown
(Source/Castro.cpp:488-505)."""
import ctypes as C

import numpy as np

from ._lib import check, lib, _dp, _ip

NVAR = 20


def _iv(v):
    return (C.c_int * 3)(*[int(a) for a in v])


class FluxRegister:
    """Register of the fine boxes `fine_boxes` (n x 6, fine index space) of a
    level refined by `ratio` from a coarse domain of `crse_n_cell` cells."""

    def __init__(self, fine_boxes, ratio, crse_n_cell, bc):
        self.L = lib()
        fb = np.ascontiguousarray(np.asarray(fine_boxes, dtype=np.int32).reshape(-1, 6))
        self.fine_boxes = fb
        self.h = self.L.mc_flux_register_create(len(fb), fb.ctypes.data_as(_ip), _iv(ratio), _iv(crse_n_cell),
                                                int(bc))
        if not self.h:
            raise RuntimeError("mc_flux_register_create: %s" % self.L.mc_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.mc_flux_register_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def data(self):
        """Every register face, in storage order (fine box, direction, side)."""
        n = self.L.mc_flux_register_size(self.h)
        out = np.zeros(n)
        check(self.L.mc_flux_register_get(self.h, out.ctypes.data_as(_dp)), "mc_flux_register_get")
        return out

    def set_val(self, v):
        check(self.L.mc_flux_register_set_val(self.h, float(v)), "mc_flux_register_set_val")

    def crse_init(self, ptr, lo, hi, d, mult):
        check(self.L.mc_flux_register_crse_init(self.h, C.cast(ptr, _dp), _iv(lo), _iv(hi), int(d), float(mult)),
              "mc_flux_register_crse_init")

    def fine_add(self, j, ptr, lo, hi, d, mult):
        check(self.L.mc_flux_register_fine_add(self.h, int(j), C.cast(ptr, _dp), _iv(lo), _iv(hi), int(d),
                                               float(mult)), "mc_flux_register_fine_add")

    def clear_internal_borders(self):
        check(self.L.mc_flux_register_clear_internal_borders(self.h), "mc_flux_register_clear_internal_borders")

    def reflux(self, state_ptr, s_lo, s_hi, lo, hi, vol_ptr, v_lo, v_hi, scale=1.0):
        check(self.L.mc_flux_register_reflux(self.h, C.cast(state_ptr, _dp), _iv(s_lo), _iv(s_hi), _iv(lo), _iv(hi),
                                             C.cast(vol_ptr, _dp), _iv(v_lo), _iv(v_hi), float(scale)),
              "mc_flux_register_reflux")


def average_down(fine_ptr, f_lo, f_hi, crse_ptr, c_lo, c_hi, lo, hi, ratio, ncomp=NVAR):
    check(lib().mc_average_down(C.cast(fine_ptr, _dp), _iv(f_lo), _iv(f_hi), C.cast(crse_ptr, _dp), _iv(c_lo),
                                _iv(c_hi), _iv(lo), _iv(hi), _iv(ratio), int(ncomp)), "mc_average_down")


def set_flux_registers(level, crse_side=None, fine_side=None):
    """The level's steps initialise crse_side (the finer level's register)
    and add to fine_side (its own)."""
    level._flux_registers = (crse_side, fine_side)  # kept alive with the level
    check(lib().mc_level_set_flux_registers(level.h, crse_side.h if crse_side else None,
                                            fine_side.h if fine_side else None), "mc_level_set_flux_registers")


def reflux(level, R):
    """Castro::reflux into the coarse level's state; zeroes the register."""
    check(lib().mc_level_reflux(level.h, R.h), "mc_level_reflux")


def average_down_levels(fine, crse, ratio, fine_box_ids=None):
    """Castro::avgDown from the given boxes of the fine level (all if None)."""
    ids = None if fine_box_ids is None else np.ascontiguousarray(fine_box_ids, dtype=np.int32)
    check(lib().mc_level_average_down(fine.h, crse.h, _iv(ratio), -1 if ids is None else len(ids),
                                      None if ids is None else ids.ctypes.data_as(_ip)), "mc_level_average_down")
