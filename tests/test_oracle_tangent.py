"""Tangent-linear build of the oracle (oracle/build/liboracle_tl.so).

The same restated arithmetic compiled with a forward-mode dual number: its
values must equal the double build bit for bit, and its derivative must
match central differences of the double build along a smooth direction.
It is the CPU tangent-linear the GPU adjoint is checked against.
"""
import numpy as np
import pytest

from minicastro import configs
from helpers import smooth_direction


@pytest.fixture(scope="module")
def problem():
    from tl_problem import OracleSedovProblem
    cfg = configs.adjoint_problem(n=12, steps=3)
    cfg["r_init"] = 0.3       # the blast covers a good part of the small grid
    cfg["shell"] = (0.2, 0.45)
    P = OracleSedovProblem(cfg)
    P.dts = P.cfl_dts(3)
    return P


def test_tangent_values_equal_double_build(problem):
    P = problem
    n = P.cfg["n_cell"]
    rng = np.random.default_rng(1)
    v = P.rho0 * smooth_direction(rng, n[0]).transpose(2, 1, 0)
    J_tl, dJ, S_tl = P.tangent(v, 0.3 * P.E, P.dts)
    J, S = P.forward(P.rho0, P.E, P.dts)
    np.testing.assert_array_equal(S_tl, S)
    assert J_tl == J


def test_tangent_matches_central_differences():
    """One step from the noisy Sedov state: the map is differentiable over the
    difference stencil along smooth directions.  (Over several steps a central
    difference can straddle a branch switch of a limiter or Riemann case and
    then differs from every one-sided derivative; the tangent-linear, like the
    adjoint, is the derivative of the branch the unperturbed run takes.)"""
    from tl_problem import OracleSedovProblem
    P = OracleSedovProblem(configs.adjoint_problem(n=16, steps=1))
    P.dts = P.cfl_dts(1)
    n = P.cfg["n_cell"]
    rng = np.random.default_rng(2)
    for q in range(3):
        v = P.rho0 * smooth_direction(rng, n[0]).transpose(2, 1, 0)
        vE = P.E * rng.standard_normal()
        _, dJ, _ = P.tangent(v, vE, P.dts)
        errs = []
        for eps in (1e-6, 1e-7):
            jp = P.forward(P.rho0 + eps * v, P.E + eps * vE, P.dts)[0]
            jm = P.forward(P.rho0 - eps * v, P.E - eps * vE, P.dts)[0]
            errs.append(abs((jp - jm) / (2 * eps) - dJ) / abs(dJ))
        assert min(errs) < 1e-6, (q, dJ, errs)


def test_tangent_is_linear(problem):
    P = problem
    n = P.cfg["n_cell"]
    rng = np.random.default_rng(3)
    v1 = rng.standard_normal(tuple(n))
    v2 = rng.standard_normal(tuple(n))
    d1 = P.tangent(v1, P.E, P.dts)[1]
    d2 = P.tangent(v2, 0.0, P.dts)[1]
    d3 = P.tangent(2.0 * v1 - 3.0 * v2, 2.0 * P.E, P.dts)[1]
    assert abs(d3 - (2.0 * d1 - 3.0 * d2)) <= 1e-10 * (abs(2.0 * d1) + abs(3.0 * d2))
