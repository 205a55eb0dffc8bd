"""Sedov problem configurations.

The reference problem (Source/Prob.F90:11-15, Source/main.cpp:79-81) is a
periodic [0, 1e9]^3 cm cube with rho = 1e4 g/cc, p = 1e21 erg/cc and a 1e52 erg
explosion inside r = 1.25e8 cm.  The workloads used here follow the
benchmark definitions: the unit cube, ambient density 1 g/cc, a low ambient
internal energy (set by the ambient pressure through the EOS, as initdata
does), the explosion energy deposited uniformly inside r = 0.01 cm with the
reference's 10^3 sub-sampling, outflow boundaries, CFL 0.5.  Energies are
chosen inside the EOS table range (1e3 - 1e13 K): ambient T ~ 1.6e5 K, blast T
~ 6e7 K.
"""
import numpy as np

# physical constants of a problem: p_ambient, dens_ambient, exp_energy, r_init, nsub
REFERENCE_PROBLEM = dict(p_ambient=1.e21, dens_ambient=1.e4, exp_energy=1.e52, r_init=1.25e8, nsub=10,
                         prob_lo=(0.0, 0.0, 0.0), prob_hi=(1.e9, 1.e9, 1.e9), bc=0)
UNIT_SEDOV = dict(p_ambient=1.e13, dens_ambient=1.0, exp_energy=1.e12, r_init=0.01, nsub=10,
                  prob_lo=(0.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 1.0), bc=1)


def grid_boxes(n_cell, max_grid):
    """AMReX-style chop of [0, n)^3 into boxes of at most max_grid cells per side."""
    n = n_cell if np.ndim(n_cell) else (n_cell,) * 3
    cuts = []
    for d in range(3):
        c = list(range(0, n[d], max_grid)) + [n[d]]
        cuts.append(c)
    boxes = []
    for k in range(len(cuts[2]) - 1):
        for j in range(len(cuts[1]) - 1):
            for i in range(len(cuts[0]) - 1):
                boxes.append([cuts[0][i], cuts[1][j], cuts[2][k],
                              cuts[0][i + 1] - 1, cuts[1][j + 1] - 1, cuts[2][k + 1] - 1])
    return np.array(boxes, dtype=np.int32)


def uneven_boxes(n=256, seed=2024, parts=4, sizes=tuple(range(16, 97, 8))):
    """Tensor partition of [0,n)^3: per axis `parts` sizes drawn uniformly from
    `sizes` (multiples of 8 in 16..96) that sum to n; parts^3 boxes."""
    rng = np.random.default_rng(seed)
    cuts = []
    for d in range(3):
        while True:
            s = rng.choice(sizes, size=parts - 1)
            last = n - int(s.sum())
            if last in sizes:
                s = list(s) + [last]
                rng.shuffle(s)
                break
        c = np.concatenate([[0], np.cumsum(s)]).astype(int)
        cuts.append(c)
    boxes = []
    for k in range(parts):
        for j in range(parts):
            for i in range(parts):
                boxes.append([cuts[0][i], cuts[1][j], cuts[2][k],
                              cuts[0][i + 1] - 1, cuts[1][j + 1] - 1, cuts[2][k + 1] - 1])
    return np.array(boxes, dtype=np.int32)


def config(idx):
    """Workload dict for benchmark configuration idx (1-based)."""
    if idx == 1:
        return dict(name="sedov32_1box", n_cell=(32,) * 3, boxes=grid_boxes(32, 32), steps=10, **UNIT_SEDOV)
    if idx == 2:
        return dict(name="sedov128_8x64", n_cell=(128,) * 3, boxes=grid_boxes(128, 64), steps=20, **UNIT_SEDOV)
    if idx == 3:
        return dict(name="sedov256_uneven64", n_cell=(256,) * 3, boxes=uneven_boxes(256), steps=10, **UNIT_SEDOV)
    raise ValueError("no configuration %r" % idx)


# Adjoint workload: 64^3 single box, 10 steps.  The deposition radius and
# energy are chosen so that the blast front reaches the objective's shell
# (r in [0.2, 0.3]) within the 10 CFL-limited steps: it starts at r = 0.15
# (9.6 cells) and moves about half a cell per step.
ADJOINT_SEDOV = dict(UNIT_SEDOV, r_init=0.15, exp_energy=1.e14)


def adjoint_problem(n=64, steps=10, seed=11):
    """Adjoint workload: Sedov on n^3 cells in one box, initial density
    rho_amb (1 + 0.01 N(0,1)) seeded, objective sum(rho^2) over cells whose
    centres lie at radius [0.2, 0.3] from the domain centre."""
    d = dict(ADJOINT_SEDOV)
    d.update(name="adjoint_sedov%d" % n, n_cell=(n,) * 3, boxes=grid_boxes(n, n), steps=steps,
             noise=0.01, seed=seed, shell=(0.2, 0.3))
    return d


def ensemble(n_inst=24, seed=7):
    """Independent single-box Sedov instances (side, E, r_dep, rho_amb seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n_inst):
        side = int(rng.choice([32, 48, 64, 96]))
        E = 10.0 ** rng.uniform(12.0, 13.0)
        r = rng.uniform(0.01, 0.04)
        rho = rng.uniform(0.5, 2.0)
        d = dict(UNIT_SEDOV)
        d.update(name="inst%02d" % m, n_cell=(side,) * 3, boxes=grid_boxes(side, side), exp_energy=E,
                 r_init=r, dens_ambient=rho, steps=50)
        out.append(d)
    return out
