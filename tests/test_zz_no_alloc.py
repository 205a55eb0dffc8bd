"""mc_calculate_blast_radius_noalloc (the allocation-free variant of
calculate_blast_radius, Source/Castro_nd.F90:349-400) against the entry point
it varies, through the C ABI: the same bits on two calls in a row, the second
on a smaller box, at the size limit, and the refusal of a box above it."""
import ctypes as C

import numpy as np
import pytest

import minicastro as mc
from helpers import Dev, table_bin

pytestmark = pytest.mark.gpu

NVAR = 20
MAX_CELLS = 4194304  # MC_NOALLOC_MAX_CELLS


@pytest.fixture(scope="module", autouse=True)
def gpu_table():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())


def sedov_state(n):
    """Density of a 3-step Sedov run on an n^3 level (GPU), Fortran order."""
    from helpers import small_cfg
    cfg = small_cfg(n=n, bc=1, r_init=0.15)
    G = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_FUSED)
    G.init_data(**cfg)
    G.run(3, use_graph=False)
    S = G.state()
    G.close()
    return S


def both(lo, hi, fab, dx, plo, phi, mx):
    """(mass, radius) of calculate_blast_radius and of the no-alloc variant,
    each accumulated from the same nonzero start values."""
    L = mc.lib()
    iv = lambda v: (C.c_int * 3)(*v)
    dv = lambda v: (C.c_double * 3)(*v)
    m0, r0 = C.c_double(1.25), C.c_double(-3.5)
    m1, r1 = C.c_double(1.25), C.c_double(-3.5)
    L.calculate_blast_radius(iv(lo), iv(hi), *fab.args(), dv(dx), dv(plo), dv(phi), C.byref(m0), C.byref(r0), mx)
    rc = L.mc_calculate_blast_radius_noalloc(iv(lo), iv(hi), *fab.args(), dv(dx), dv(plo), dv(phi), C.byref(m1),
                                             C.byref(r1), mx)
    assert rc == 0, L.mc_last_error()
    return (m0.value, r0.value), (m1.value, r1.value)


def test_noalloc_equals_calculate_blast_radius_two_calls():
    S = sedov_state(64)
    mx = float(S[..., 0].max())
    dx = (1.0 / 64,) * 3
    # first call: the whole 64^3 level as one box (1024 blocks)
    fab = Dev((0, 0, 0), (63, 63, 63), NVAR, S)
    ref, got = both((0, 0, 0), (63, 63, 63), fab, dx, (0, 0, 0), (1, 1, 1), mx)
    assert got == ref
    assert ref[0] > 1.25
    # second call, smaller: a ragged sub-box of a FAB with ghost cells (a few
    # blocks, so the first call's partials beyond them must not be read)
    lo, hi = (20, 22, 18), (43, 41, 33)
    flo, fhi = (18, 20, 16), (45, 43, 35)
    sub = np.asfortranarray(S[flo[0]:fhi[0] + 1, flo[1]:fhi[1] + 1, flo[2]:fhi[2] + 1])
    fab2 = Dev(flo, fhi, NVAR, sub)
    ref2, got2 = both(lo, hi, fab2, dx, (0, 0, 0), (1, 1, 1), mx)
    assert got2 == ref2
    assert ref2[0] > 1.25 and ref2 != ref


def test_noalloc_at_the_size_limit_and_refused_above():
    import torch
    L = mc.lib()
    shape = (256, 128, 128)  # MAX_CELLS cells
    assert int(np.prod(shape)) == MAX_CELLS
    rng = np.random.default_rng(3)
    rho = 1.0 + 0.2 * rng.random(shape)
    S = np.zeros(shape + (NVAR,), order="F")
    S[..., 0] = rho
    fab = Dev((0, 0, 0), tuple(s - 1 for s in shape), NVAR, S)
    dx = (1.0 / 256,) * 3
    ref, got = both((0, 0, 0), tuple(s - 1 for s in shape), fab, dx, (0, 0, 0), (1, 0.5, 0.5), 1.15)
    assert got == ref and ref[0] > 1.25
    # one plane more than the limit: refused before anything is launched
    m, r = C.c_double(0.0), C.c_double(0.0)
    iv = lambda v: (C.c_int * 3)(*v)
    dv = lambda v: (C.c_double * 3)(*v)
    rc = L.mc_calculate_blast_radius_noalloc(iv((0, 0, 0)), iv((255, 127, 128)), *fab.args(), dv(dx),
                                             dv((0, 0, 0)), dv((1, 1, 1)), C.byref(m), C.byref(r), 1.15)
    assert rc == -1 and b"exceeds the limit" in L.mc_last_error()
    assert m.value == 0.0 and r.value == 0.0
    torch.cuda.synchronize()
    # and the entry point still works after the refusal
    ref3, got3 = both((0, 0, 0), (31, 31, 31), fab, dx, (0, 0, 0), (1, 0.5, 0.5), 1.15)
    assert got3 == ref3
