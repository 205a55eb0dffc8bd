"""Op-count build of the oracle (oracle/build/liboracle_cnt.so).

The same restated arithmetic with a scalar that counts its floating-point
operations; it gives the FP64 work per zone-update of the reference
algorithm.  Its values must equal the double build bit for bit, its counts
must be reproducible, and they must follow the structure of the work: a
pointwise kernel's count is linear in the number of cells, and splitting a
domain into more boxes adds the ghost-cell work the reference repeats per
box.
"""
import numpy as np
import pytest

import orc
from helpers import oracle_table, small_cfg, table_array
from opcount_zone_update import count_steps


@pytest.fixture(scope="module", autouse=True)
def tables():
    oracle_table()
    orc.cnt_eos_init_arrays(table_array())


def test_counted_steps_equal_double_build():
    cfg = small_cfg(n=16, r_init=0.2)
    res = count_steps(cfg, skip=1, count=2, nthreads=4)
    for dt, zones, ops, same in res:
        assert same
        assert zones == 16 ** 3
        assert all(ops[k] > 0 for k in orc.OPCOUNT_FIELDS)


def test_counts_reproducible():
    cfg = small_cfg(n=12)
    a = count_steps(cfg, skip=0, count=1)[0][2]
    b = count_steps(cfg, skip=0, count=1)[0][2]
    assert a == b


def _compute_temp_ops(ncell, rho, T):
    """ops of compute_temp (Source/Castro_nd.F90:253-297) on ncell copies of
    one cleaned cell."""
    L = orc.cnt_lib()
    u = np.zeros((ncell, 1, 1, orc.NVAR), order="F")
    u[..., 0] = rho
    u[..., orc.NVAR - 13:] = rho / 13.0
    e = 1.0e15
    u[..., 4] = rho * e
    u[..., 5] = rho * e
    u[..., 6] = T
    lo, hi = orc.iv(0, 0, 0), orc.iv(ncell - 1, 0, 0)
    L.orc_opcount_reset()
    L.orc_compute_temp(lo, hi, orc.ptr(u), lo, hi)
    out = (orc.C.c_ulonglong * 5)()
    L.orc_opcount_read(out)
    return list(out), u


def test_pointwise_count_is_linear_in_cells():
    one, u1 = _compute_temp_ops(1, 2.0, 1.0e7)
    many, u7 = _compute_temp_ops(7, 2.0, 1.0e7)
    assert one[0] > 0 and one[4] > 0          # adds and at least one log
    assert many == [7 * v for v in one]
    np.testing.assert_array_equal(u7[:1], u1)


def test_more_boxes_count_more_ghost_work():
    from minicastro import configs
    one = small_cfg(n=16)
    eight = dict(one, boxes=configs.grid_boxes(16, 8))
    c1 = count_steps(one, skip=0, count=1)[0]
    c8 = count_steps(eight, skip=0, count=1)[0]
    assert c1[3] and c8[3]
    assert c1[0] == c8[0]                      # same dt: the result does not depend on the boxes
    f1 = sum(c1[2][k] for k in ("add", "mul", "div", "sqrt"))
    f8 = sum(c8[2][k] for k in ("add", "mul", "div", "sqrt"))
    # eight 8^3 boxes grow to 16^3 with ghosts against 24^3 for one box; only
    # part of the step runs on the grown boxes
    assert f8 > 1.15 * f1
