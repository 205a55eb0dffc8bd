"""Next-row kernels against the oracle: the blast-radius diagnostic of
post_timestep (Source/Castro.cpp:371-424 with Source/Castro_nd.F90:349-400),
density-gradient tagging (Source/Castro_nd.F90:301-345) and GPU initdata
(Source/Prob.F90:19-129, compared bit for bit)."""
import ctypes as C

import numpy as np
import pytest

import orc
import minicastro as mc
from minicastro import configs
from helpers import Dev, gpu_call, oracle_level, oracle_table, small_cfg, table_bin

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu_table():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())
    oracle_table()


@pytest.mark.parametrize("mode", [mc.MODE_PLAIN, mc.MODE_FUSED])
def test_blast_radius_matches_oracle(mode):
    cfg = small_cfg(n=24, bc=1, r_init=0.15)
    boxes = configs.grid_boxes(24, 12)
    O = oracle_level(cfg, boxes=boxes)
    G = mc.Level(cfg["n_cell"], boxes, cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mode)
    G.init_data(**cfg)
    for _ in range(5):
        O.coarse_step()
        G.coarse_step()
    S = O.state()
    mx = S[..., 0].max()
    m, r = C.c_double(0.0), C.c_double(0.0)
    dx = (1.0 / 24,) * 3
    for b in boxes:
        lo, hi = tuple(b[:3]), tuple(b[3:])
        sub = np.asfortranarray(S[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1])
        orc.call("calculate_blast_radius", orc.iv(*lo), orc.iv(*hi), orc.Fab(lo, hi, 20, sub), orc.dv(*dx),
                 orc.dv(0, 0, 0), orc.dv(1, 1, 1), C.byref(m), C.byref(r), float(mx))
    gm, gr, gx = G.blast_radius()
    assert gx == mx
    assert gm == pytest.approx(m.value, rel=1e-14) and gr == pytest.approx(r.value, rel=1e-13)
    assert gm > 0 and 0.0 < gr / gm < 0.5


def test_denerror_tags_bitwise():
    rng = np.random.default_rng(9)
    lo, hi = (2, 3, 1), (9, 8, 7)
    dlo, dhi = (1, 2, 0), (10, 9, 8)
    den = 1.0 + rng.random(tuple(h - l + 1 for l, h in zip(dlo, dhi))) * rng.choice([0.01, 1.0], 1)[0]
    den[4:6, 3:5, 2:4] += 2.0
    tshape = tuple(h - l + 1 for l, h in zip(lo, hi))
    to = np.zeros(tshape, dtype=np.int8, order="F")
    oden = orc.Fab(dlo, dhi, 1, den)
    orc.lib().orc_denerror(orc.iv(*lo), orc.iv(*hi), to.ctypes.data_as(C.c_void_p), orc.iv(*lo), orc.iv(*hi),
                           oden.args()[0], orc.iv(*dlo), orc.iv(*dhi), C.c_byte(1), C.c_byte(0))
    import torch
    tg = torch.zeros(int(np.prod(tshape)), dtype=torch.int8, device="cuda")
    gd = Dev(dlo, dhi, 1, den)
    mc.lib().denerror(orc.iv(*lo), orc.iv(*hi), C.c_void_p(tg.data_ptr()), orc.iv(*lo), orc.iv(*hi), *gd.args(),
                      1, 0)
    torch.cuda.synchronize()
    got = tg.cpu().numpy().reshape(tshape, order="F")
    np.testing.assert_array_equal(got, to)
    assert 0 < got.sum() < got.size


def test_initdata_reference_problem_bitwise():
    """The reference's own problem constants (periodic 1e9 cm cube, 1e52 erg)."""
    p = configs.REFERENCE_PROBLEM
    cfg = dict(p, n_cell=(16, 16, 16), boxes=configs.grid_boxes(16, 8))
    O = oracle_level(cfg)
    G = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_PLAIN)
    G.init_data(**cfg)
    np.testing.assert_array_equal(G.state(), O.state())
    for _ in range(2):
        assert O.coarse_step() == G.coarse_step()
    np.testing.assert_array_equal(G.state(), O.state())
