"""Shared test helpers: table loading, oracle states, device buffers."""
import ctypes as C
import os

import numpy as np

import orc
import minicastro as mc
from minicastro import configs

NVAR, QVAR, NQAUX, NGDNV, NSPEC = 20, 21, 4, 6, 13


def table_bin():
    p = mc.table_path("bin")
    if not os.path.exists(p):
        mc.build()
    return p


def table_array():
    """(21, 201, 541) table from the generated binary cache."""
    raw = np.fromfile(table_bin(), dtype=np.uint8)
    assert raw[:8].tobytes() == b"MCHELM01"
    dims = np.frombuffer(raw[8:16].tobytes(), dtype=np.int32)
    assert tuple(dims) == (541, 201)
    return np.frombuffer(raw[16:].tobytes(), dtype=np.float64).reshape(21, 201, 541)


_oracle_table = False


def oracle_table():
    global _oracle_table
    if not _oracle_table:
        orc.eos_init_arrays(table_array())
        _oracle_table = True


def oracle_level(cfg, nthreads=1, boxes=None):
    oracle_table()
    b = cfg["boxes"] if boxes is None else boxes
    L = orc.Level(cfg["n_cell"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], b, nthreads)
    L.init_data(cfg["p_ambient"], cfg["dens_ambient"], cfg["exp_energy"], cfg["r_init"], cfg["nsub"])
    return L


def small_cfg(n=16, bc=1, r_init=0.12, **kw):
    d = dict(configs.UNIT_SEDOV)
    d.update(name="small", n_cell=(n, n, n), boxes=configs.grid_boxes(n, n), steps=3, bc=bc, r_init=r_init)
    d.update(kw)
    return d


def smooth_direction(rng, n, modes=6, kmax=3):
    """Random smooth field on an n^3 grid (z, y, x order): a sum of `modes`
    separable cosines with wavenumbers <= kmax, random phases and Gaussian
    amplitudes."""
    x = (np.arange(n) + 0.5) / n
    v = np.zeros((n, n, n))
    for _ in range(modes):
        k = rng.integers(0, kmax + 1, 3)
        ph = rng.uniform(0.0, 2.0 * np.pi, 3)
        a = rng.standard_normal()
        v += (a * np.cos(2 * np.pi * k[2] * x + ph[2])[:, None, None] *
              np.cos(2 * np.pi * k[1] * x + ph[1])[None, :, None] *
              np.cos(2 * np.pi * k[0] * x + ph[0])[None, None, :])
    return v


def rel_err(a, b):
    """max-norm relative difference per component (last axis)."""
    a = np.asarray(a).reshape(-1, a.shape[-1])
    b = np.asarray(b).reshape(-1, b.shape[-1])
    den = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return np.abs(a - b).max(axis=0) / den


# ------------------------------------------------------------ device buffers
def torch_cuda():
    import torch
    return torch


class Dev:
    """A device copy of a Fortran-ordered numpy FAB with its (lo, hi)."""

    def __init__(self, lo, hi, nc, host=None):
        torch = torch_cuda()
        self.lo, self.hi, self.nc = tuple(lo), tuple(hi), nc
        self.shape = tuple(h - l + 1 for l, h in zip(lo, hi)) + (nc,)
        n = int(np.prod(self.shape))
        if host is None:
            self.t = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        else:
            h = np.asarray(host, dtype=np.float64).reshape(self.shape, order="F")
            self.t = torch.from_numpy(h.ravel(order="F").copy()).to("cuda")

    def args(self):
        return (C.cast(self.t.data_ptr(), C.POINTER(C.c_double)), (C.c_int * 3)(*self.lo), (C.c_int * 3)(*self.hi))

    def host(self):
        import torch
        torch.cuda.synchronize()
        return self.t.cpu().numpy().reshape(self.shape, order="F")


def gpu_call(name, *args):
    fl = []
    for a in args:
        if isinstance(a, Dev):
            fl.extend(a.args())
        else:
            fl.append(a)
    getattr(mc.lib(), name)(*fl)
