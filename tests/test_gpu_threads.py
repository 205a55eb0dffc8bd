"""Re-entrancy of the drop-in ABI across host threads.

The reference host calls the per-box entry points from several OpenMP
threads at once (Source/Castro_hydro.cpp:28-31), so they must be re-entrant
and thread-safe: the stream and the error state are per host thread, the
EOS table is shared read-only.  Four per-box level drivers, each with its
own stream and problem, step concurrently from four host threads (ctypes
releases the GIL during each call); every state and dt must equal the same
run made alone, bit for bit.
"""
import threading

import numpy as np
import pytest

import minicastro as mc
from minicastro import configs
from helpers import small_cfg, table_bin

pytestmark = pytest.mark.gpu

CASES = [
    dict(n=24, bc=1, r_init=0.10, max_grid=12),
    dict(n=16, bc=0, r_init=0.15, max_grid=8),
    dict(n=20, bc=1, r_init=0.20, max_grid=10),
    dict(n=32, bc=0, r_init=0.08, max_grid=16),
]


@pytest.fixture(scope="module", autouse=True)
def gpu_table():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())


def run(case, steps=3):
    cfg = small_cfg(n=case["n"], bc=case["bc"], r_init=case["r_init"])
    boxes = configs.grid_boxes(case["n"], case["max_grid"])
    L = mc.Level(cfg["n_cell"], boxes, cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_PLAIN)
    L.init_data(**cfg)
    dts = [L.coarse_step() for _ in range(steps)]
    S = L.state()
    L.close()
    return dts, S


def test_concurrent_levels_equal_sequential():
    alone = [run(c) for c in CASES]
    out = [None] * len(CASES)
    err = []

    def work(q):
        try:
            out[q] = run(CASES[q])
        except Exception as e:  # reported below
            err.append((q, repr(e)))

    th = [threading.Thread(target=work, args=(q,)) for q in range(len(CASES))]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not err, err
    assert all(not t.is_alive() for t in th)
    for q, ((dta, Sa), (dtc, Sc)) in enumerate(zip(alone, out)):
        assert dta == dtc, q
        assert np.array_equal(Sa.view(np.uint64), Sc.view(np.uint64)), q
