"""The C ABI library loads without a GPU and exports every entry point that
include/mc_castro.h declares (the names a mini-Castro host binds,
Source/Castro_F.H:11-161, plus the mc_ extensions).  No compute calls."""
import ctypes as C
import os
import re

import pytest

import minicastro as mc

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mc_castro.h")
REFERENCE_NAMES = ["network_init", "network_finalize", "eos_init", "eos_finalize", "ctoprim",
                   "enforce_minimum_density", "normalize_species", "reset_internal_e", "compute_temp", "estdt",
                   "store_flux", "fill_hydro_source", "normalize_species_fluxes", "apply_av", "trans1", "trans2",
                   "compute_flux", "divu", "trace_ppm", "initdata", "denerror", "calculate_blast_radius"]


def declared():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof")))


def test_header_declares_reference_entry_points():
    d = declared()
    for n in REFERENCE_NAMES:
        assert n in d, n


def test_library_exports_every_declared_symbol():
    L = C.CDLL(mc.lib_path())
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    bound = {n for n, _, _ in mc._lib.SIGS}
    assert set(declared()) <= bound, set(declared()) - bound


def test_level_create_reports_instead_of_aborting():
    """Without a GPU runtime mc_level_create returns NULL with the HIP error
    recorded for mc_last_error; with one it returns a level."""
    import torch
    L = mc.lib()
    lvl = L.mc_level_create((C.c_int * 3)(8, 8, 8), (C.c_double * 3)(0, 0, 0), (C.c_double * 3)(1, 1, 1), 1, 1,
                            (C.c_int * 6)(0, 0, 0, 7, 7, 7))
    if torch.cuda.is_available():
        assert lvl
        L.mc_level_destroy(lvl)
    else:
        assert not lvl
        assert L.mc_last_error()


def test_bad_box_lists_rejected():
    L = mc.lib()
    mk = lambda boxes: L.mc_level_create((C.c_int * 3)(8, 8, 8), (C.c_double * 3)(0, 0, 0),
                                         (C.c_double * 3)(1, 1, 1), 1, len(boxes) // 6, (C.c_int * len(boxes))(*boxes))
    assert not mk([0, 0, 0, 8, 7, 7])            # outside the domain
    assert not mk([0, 0, 0, 3, 7, 7])            # does not tile the domain
    assert not mk([0, 0, 0, 7, 7, 7, 0, 0, 0, 7, 7, 7])  # overlap / double count
    assert not mk([0, 0, 0, 7, 7, 3, 0, 0, 0, 7, 7, 3])  # overlap, volumes add up to the domain
    assert b"tile" in L.mc_last_error() or b"domain" in L.mc_last_error() or b"overlap" in L.mc_last_error()



def test_noalloc_blast_radius_refuses_an_oversized_box_before_any_launch():
    """MC_NOALLOC_MAX_CELLS cells at most; the refusal comes before any
    device call, so it runs without a GPU."""
    L = mc.lib()
    src = open(HDR).read()
    cap = int(re.search(r"#define MC_NOALLOC_MAX_CELLS (\d+)", src).group(1))
    m, r = C.c_double(0.0), C.c_double(0.0)
    iv = lambda v: (C.c_int * 3)(*v)
    dv = lambda v: (C.c_double * 3)(*v)
    hi = (cap // (128 * 128), 127, 127)   # one plane more than the cap
    rc = L.mc_calculate_blast_radius_noalloc(iv((0, 0, 0)), iv(hi), None, iv((0, 0, 0)), iv(hi), dv((1, 1, 1)),
                                             dv((0, 0, 0)), dv((1, 1, 1)), C.byref(m), C.byref(r), 1.0)
    assert rc == -1 and b"exceeds the limit" in L.mc_last_error()
    assert m.value == 0.0 and r.value == 0.0
