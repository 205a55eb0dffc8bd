"""Discrete adjoint of the CTU step (mc_adjoint_*).

Arrays are device buffers given by address (for instance torch.Tensor
.data_ptr() of a contiguous float64 CUDA tensor) holding NVAR x cells in
Fortran order: component slowest, i fastest -- a tensor of shape
(NVAR, nz, ny, nx).
"""
import ctypes as C

import numpy as np

from ._lib import lib, check, _dp

NVAR = 20


class Adjoint:
    def __init__(self, n_cell, prob_lo, prob_hi, bc, max_steps):
        self.L = lib()
        nc = (C.c_int * 3)(*[int(v) for v in n_cell])
        lo = (C.c_double * 3)(*[float(v) for v in prob_lo])
        hi = (C.c_double * 3)(*[float(v) for v in prob_hi])
        self.n_cell = tuple(int(v) for v in n_cell)
        self.max_steps = int(max_steps)
        self.h = self.L.mc_adjoint_create(nc, lo, hi, int(bc), self.max_steps)
        if not self.h:
            raise RuntimeError("mc_adjoint_create: %s" % self.L.mc_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.mc_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, u0_ptr, uout_ptr, nsteps, dts=None):
        """Advance nsteps from the state at u0_ptr into uout_ptr, keeping
        checkpoints.  dts=None: CFL-limited dt sequence, returned; else the
        given sequence is used.  Returns the dt sequence (numpy)."""
        compute = dts is None
        d = np.zeros(nsteps) if compute else np.ascontiguousarray(dts, dtype=np.float64).copy()
        check(self.L.mc_adjoint_forward(self.h, C.c_void_p(u0_ptr), C.c_void_p(uout_ptr), int(nsteps),
                                        d.ctypes.data_as(_dp), 1 if compute else 0), "mc_adjoint_forward")
        return d

    def backward(self, ub_final_ptr, ub0_ptr):
        """Adjoint of the initial state of the last forward pass, given the
        adjoint of its final state."""
        check(self.L.mc_adjoint_backward(self.h, C.c_void_p(ub_final_ptr), C.c_void_p(ub0_ptr)),
              "mc_adjoint_backward")

    @property
    def keeps_intermediates(self):
        """True when the forward pass keeps every step's intermediates (the
        reverse pass then reads them instead of recomputing each step)."""
        return bool(self.L.mc_adjoint_keeps_intermediates(self.h))

    def keep_intermediates(self, keep):
        """Select keeping (True) or recomputing (False) the intermediates;
        drops the last forward pass."""
        check(self.L.mc_adjoint_keep_intermediates(self.h, 1 if keep else 0), "mc_adjoint_keep_intermediates")
