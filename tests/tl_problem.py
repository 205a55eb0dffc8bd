"""Sedov adjoint workload on the CPU oracle: the initial state as a function
of the density field and the explosion energy (numpy), and the objective."""
import numpy as np

import orc
from helpers import oracle_level, oracle_table, table_array

NVAR, URHO, UEDEN, UEINT, UTEMP, UFS = 20, 0, 4, 5, 6, 7


class OracleSedovProblem:
    """U0(rho0, E): initdata (Source/Prob.F90:19-129) with a per-cell density
    field; shapes are (nx, ny, nz[, NVAR]) in Fortran order."""

    def __init__(self, cfg):
        oracle_table()
        self.cfg = cfg
        E = cfg["exp_energy"]
        rho_amb = cfg["dens_ambient"]
        c0 = dict(cfg, exp_energy=0.0)
        S1 = oracle_level(cfg).state()
        S0 = oracle_level(c0).state()
        self.a = S0[..., UEINT] / rho_amb
        self.b = (S1[..., UEINT] - S0[..., UEINT]) / rho_amb / E
        self.T0 = S1[..., UTEMP].copy()
        n = cfg["n_cell"]
        rng = np.random.default_rng(cfg.get("seed", 11))
        z = rng.standard_normal((n[2], n[1], n[0])).transpose(2, 1, 0)  # same draws as the GPU harness
        self.rho0 = rho_amb * (1.0 + cfg.get("noise", 0.01) * z)
        self.E = float(E)
        lo, hi = np.asarray(cfg["prob_lo"]), np.asarray(cfg["prob_hi"])
        dx = (hi - lo) / np.asarray(n)
        ctr = 0.5 * (lo + hi)
        xs = [lo[d] + dx[d] * (np.arange(n[d]) + 0.5) - ctr[d] for d in range(3)]
        r = np.sqrt(xs[0][:, None, None] ** 2 + xs[1][None, :, None] ** 2 + xs[2][None, None, :] ** 2)
        r0, r1 = cfg.get("shell", (0.2, 0.3))
        self.mask = ((r >= r0) & (r <= r1)).astype(np.float64)

    def U0(self, rho0, E, drho0=None, dE=0.0):
        n = self.cfg["n_cell"]
        U = np.zeros(tuple(n) + (NVAR,), order="F")
        dU = np.zeros_like(U)
        eint = self.a + self.b * E
        U[..., URHO] = rho0
        U[..., UEDEN] = rho0 * eint
        U[..., UEINT] = rho0 * eint
        U[..., UTEMP] = self.T0
        U[..., UFS] = rho0
        if drho0 is not None:
            deint = self.b * dE
            dU[..., URHO] = drho0
            dU[..., UEDEN] = drho0 * eint + rho0 * deint
            dU[..., UEINT] = drho0 * eint + rho0 * deint
            dU[..., UFS] = drho0
        return U, dU

    def objective(self, U):
        return float((self.mask * U[..., URHO] ** 2).sum())

    def forward(self, rho0, E, dts, nthreads=8):
        """J of the plain oracle with the frozen dt sequence."""
        cfg = self.cfg
        L = orc.Level(cfg["n_cell"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], cfg["boxes"], nthreads)
        L.set_state(self.U0(rho0, E)[0])
        for dt in dts:
            L.advance(dt)
            L.post_timestep()
        return self.objective(L.state()), L.state()

    def tangent(self, drho0, dE, dts, nthreads=8):
        """(J, dJ) along (drho0, dE) by the tangent-linear oracle."""
        cfg = self.cfg
        orc.tl_eos_init_arrays(table_array())
        L = orc.TLLevel(cfg["n_cell"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], cfg["boxes"], nthreads)
        U, dU = self.U0(self.rho0, self.E, drho0, dE)
        L.set_state(U, dU)
        for dt in dts:
            L.step(dt)
        S, dS = L.state()
        return self.objective(S), float((self.mask * 2.0 * S[..., URHO] * dS[..., URHO]).sum()), S

    def cfl_dts(self, nsteps, nthreads=8):
        """dt_0 = 0.5 estdt(U0), dt_{n+1} = min(0.5 estdt(U_{n+1}), 1.1 dt_n)."""
        cfg = self.cfg
        L = orc.Level(cfg["n_cell"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], cfg["boxes"], nthreads)
        L.set_state(self.U0(self.rho0, self.E)[0])
        dts = []
        for n in range(nsteps):
            est = L.est_time_step()
            dt = est if n == 0 else min(est, 1.1 * dts[-1])
            dts.append(dt)
            L.advance(dt)
            L.post_timestep()
        return np.array(dts)
