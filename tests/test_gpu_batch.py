"""mc_advance_batch: Castro::advance + post_timestep over caller-owned FABs.

Every box of the level advances in one launch per stage while the caller
keeps its per-box FABs.  Each step's S_new and the returned estTimeStep must
equal the per-box level driver (the reference kernel order) bit for bit.
The S_old FABs carry NaN ghost cells, which must never be read; S_new FABs
of a different extent keep their ghost cells untouched; S_new may alias
S_old.
"""
import numpy as np
import pytest

import minicastro as mc
from minicastro import configs
from helpers import Dev, small_cfg, table_bin

pytestmark = pytest.mark.gpu

NVAR = 20


@pytest.fixture(scope="module", autouse=True)
def gpu_table():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())
    yield
    mc.batch.release()


def uneven(n, cuts):
    """Tensor-product box list of [0,n)^3 with the given interior cuts."""
    c = [0] + list(cuts) + [n]
    return [[c[i], c[j], c[k], c[i + 1] - 1, c[j + 1] - 1, c[k + 1] - 1]
            for k in range(len(c) - 1) for j in range(len(c) - 1) for i in range(len(c) - 1)]


def fabs_for(boxes, S, ghost, fill_ghost=float("nan")):
    """Device FABs of the boxes grown by `ghost`, valid cells from S."""
    out = []
    for b in boxes:
        lo = np.array(b[:3]) - ghost
        hi = np.array(b[3:]) + ghost
        shape = tuple(hi - lo + 1) + (NVAR,)
        h = np.full(shape, fill_ghost, order="F")
        g = ghost
        h[g:shape[0] - g, g:shape[1] - g, g:shape[2] - g, :] = S[b[0]:b[3] + 1, b[1]:b[4] + 1, b[2]:b[5] + 1, :]
        out.append(Dev(lo, hi, NVAR, h))
    return out


def valid(fab, b, ghost):
    h = fab.host()
    g = ghost
    return h[g:h.shape[0] - g, g:h.shape[1] - g, g:h.shape[2] - g, :]


def run(cfg, boxes, steps, new_ghost=2, in_place=False):
    L = mc.Level(cfg["n_cell"], boxes, cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_PLAIN)
    L.init_data(**cfg)
    S = L.state()
    old = fabs_for(boxes, S, 4)
    new = old if in_place else fabs_for(boxes, np.zeros_like(S), new_ghost, fill_ghost=-7.0)
    ng = 4 if in_place else new_ghost
    dt = L.initial_dt()
    for s in range(steps):
        work = [mc.work_item(b[:3], b[3:], o.t.data_ptr(), o.lo, o.hi, w.t.data_ptr(), w.lo, w.hi)
                for b, o, w in zip(boxes, old, new)]
        print("step", s, "dt", dt, flush=True)
        est = mc.advance_batch(cfg["n_cell"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], work, dt)
        L.advance(dt)
        L.post_timestep()
        est_ref = L.est_time_step()
        Sref = L.state()
        for b, w in zip(boxes, new):
            ref = Sref[b[0]:b[3] + 1, b[1]:b[4] + 1, b[2]:b[5] + 1, :]
            got = valid(w, b, ng)
            assert np.array_equal(got.view(np.uint64), np.ascontiguousarray(ref).view(np.uint64)), (s, b)
            if not in_place:
                h = w.host()
                inner = np.zeros(h.shape[:3], bool)
                inner[ng:h.shape[0] - ng, ng:h.shape[1] - ng, ng:h.shape[2] - ng] = True
                assert np.all(h[~inner] == -7.0)        # ghost cells of S_new untouched
        assert est == est_ref, (s, est, est_ref)
        if not in_place:
            # the next step starts from this S_new: copy it into the S_old FABs
            old = fabs_for(boxes, Sref, 4)
        dt = min(est, 1.1 * dt)


def test_batch_equals_per_box_uneven_outflow():
    cfg = small_cfg(n=40, r_init=0.1)
    run(cfg, uneven(40, [12, 28]), steps=3)


def test_batch_equals_per_box_periodic():
    cfg = small_cfg(n=32, bc=0, r_init=0.15)
    run(cfg, uneven(32, [8, 20]), steps=2)


def test_batch_in_place():
    cfg = small_cfg(n=32, r_init=0.1)
    run(cfg, configs.grid_boxes(32, 16), steps=2, in_place=True)


def test_batch_rejects_bad_work_lists():
    cfg = small_cfg(n=16)
    boxes = configs.grid_boxes(16, 8)
    S = np.zeros((16, 16, 16, NVAR), order="F")
    f = fabs_for(boxes, S, 0, fill_ghost=0.0)
    mk = lambda bl, fl: [mc.work_item(b[:3], b[3:], o.t.data_ptr(), o.lo, o.hi, o.t.data_ptr(), o.lo, o.hi)
                         for b, o in zip(bl, fl)]
    # overlapping boxes whose volumes add up to the domain
    bad = [list(b) for b in boxes]
    bad[1] = list(bad[0])
    with pytest.raises(RuntimeError, match="tile"):
        mc.advance_batch(cfg["n_cell"], cfg["prob_lo"], cfg["prob_hi"], 1, mk(bad, fabs_for(bad, S, 0, 0.0)), 1e-12)
    # a FAB that does not contain its box
    small = [Dev(np.array(b[:3]) + 1, b[3:], NVAR, None) for b in boxes]
    with pytest.raises(RuntimeError, match="contain"):
        mc.advance_batch(cfg["n_cell"], cfg["prob_lo"], cfg["prob_hi"], 1, mk(boxes, small), 1e-12)
