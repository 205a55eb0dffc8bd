"""Batched advance over caller-owned FABs (``mc_advance_batch``).

One call is Castro::advance followed by post_timestep's clean_state
(Source/Castro_advance.cpp:11-100, Source/Castro.cpp:343-369) for every box
of a single level, with one launch per stage over all boxes.  The caller's
device FABs stay where they are: each work item names a valid box and the
(pointer, lo, hi) of the S_old and S_new FABs that hold it (Fortran order,
20 components).
"""
import ctypes as C

from ._lib import BatchBox, check, lib


def _iv(v):
    return (C.c_int * 3)(*[int(a) for a in v])


def work_item(lo, hi, s_old, old_lo, old_hi, s_new, new_lo, new_hi):
    """s_old / s_new: device pointers (int) of the FABs."""
    return BatchBox(_iv(lo), _iv(hi), C.c_void_p(int(s_old)), _iv(old_lo), _iv(old_hi),
                    C.c_void_p(int(s_new)), _iv(new_lo), _iv(new_hi))


def advance_batch(n_cell, prob_lo, prob_hi, bc, work, dt):
    """Advance every box of `work` by dt; returns Castro::estTimeStep of the
    new state (cfl * min dx/(c+|u|))."""
    arr = (BatchBox * len(work))(*work)
    est = C.c_double(0.0)
    check(lib().mc_advance_batch(_iv(n_cell), (C.c_double * 3)(*prob_lo), (C.c_double * 3)(*prob_hi), int(bc),
                                 len(work), arr, float(dt), C.byref(est)), "mc_advance_batch")
    return est.value


def release():
    """Free the engine kept for the last geometry and box list."""
    lib().mc_advance_batch_release()
