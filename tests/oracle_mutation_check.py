"""Mutation check of the oracle pins (test infrastructure; not collected by
pytest).  Builds copies of the oracle (oracle/*.c) with one plausible slip
each -- a flipped sign, a wrong wave, a dropped term, a wrong index -- and
runs the CPU pin tests against every copy through ORC_LIB.  Every slip must
make a test fail, except M1: the plus state's reference state is the u-c
wave's own integral (Source/ppm.F90:529-537), so `dum` there is exactly zero
and flipping its sign changes no result.

    python tests/oracle_mutation_check.py [test files...]
"""
import os
import shutil
import subprocess
import sys
import tempfile
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
M = [
 ("M1 plus alpham -dum -> +dum (always zero: equivalent)",
  "orc_hydro.c", 0, "real alpham = HALF * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;",
  "real alpham = HALF * (dptotm * rho_ref_inv * cc_ref_inv + dum) * rho_ref * cc_ref_inv;"),
 ("M1b minus alpham -dum -> +dum", "orc_hydro.c", 1,
  "real alpham = HALF * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;",
  "real alpham = HALF * (dptotm * rho_ref_inv * cc_ref_inv + dum) * rho_ref * cc_ref_inv;"),
 ("M1c plus alphap +dup -> -dup", "orc_hydro.c", 0,
  "real alphap = HALF * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;",
  "real alphap = HALF * (dptotp * rho_ref_inv * cc_ref_inv - dup) * rho_ref * cc_ref_inv;"),
 ("M2 trans1 (gamc-1) -> (gamc+1)", "orc_hydro.c", 0, "pav * du * (gamc - ONE)", "pav * du * (gamc + ONE)"),
 ("M2b trans2 (gamc-1) -> (gamc+1)", "orc_hydro.c", 0,
  "p2av * du2 * (A(qaux, il1, jl1, kl1, QGAMC) - ONE)", "p2av * du2 * (A(qaux, il1, jl1, kl1, QGAMC) + ONE)"),
 ("M3 -pdivu -> +pdivu", "orc_hydro.c", 0, "A(src, i, j, k, n) = A(src, i, j, k, n) - pdivu;",
  "A(src, i, j, k, n) = A(src, i, j, k, n) + pdivu;"),
 ("M4 species flux normalisation removed", "orc_hydro.c", 0,
  "for (int n = UFS; n < UFS + ORC_NSPEC; ++n) A(flux, i, j, k, n) = A(flux, i, j, k, n) * fac;", "(void)fac;"),
 ("M5 species traced with u-c", "orc_hydro.c", 0, "if (do_p) A(qp, i, j, k, n) = Im_sp[1];",
  "if (do_p) A(qp, i, j, k, n) = Im_sp[0];"),
 ("M6 trans1 species flux sign", "orc_hydro.c", 0,
  "real compu = lq2[QRHO] * lq2[nqp] - cdtdx * (A(f1, ir, jr, kr, n) - A(f1, il, jl, kl, n));",
  "real compu = lq2[QRHO] * lq2[nqp] + cdtdx * (A(f1, ir, jr, kr, n) - A(f1, il, jl, kl, n));"),
 ("M7 minus alpha0r not negated", "orc_hydro.c", 0, "alpha0r = (un > ZERO) ? -alpha0r : ZERO;",
  "alpha0r = (un > ZERO) ? alpha0r : ZERO;"),
 ("M8 trans2 dt/dx of flux 3 taken from direction 2", "orc_hydro.c", 1,
  "cdtdx3 * (A(f3, ir3, jr3, kr3, URHO)", "cdtdx2 * (A(f3, ir3, jr3, kr3, URHO)"),
 ("M9 pdivu uses dx[0] in y", "orc_hydro.c", 0, "(A(qy, i, j + 1, k, GDV) - A(qy, i, j, k, GDV)) / dx[1]",
  "(A(qy, i, j + 1, k, GDV) - A(qy, i, j, k, GDV)) / dx[0]"),
 ("M10 trace minus state: transverse velocity from the u+c wave", "orc_hydro.c", 0,
  "A(qm, ii, jj, kk, QUT) = Ip[1][QUT];", "A(qm, ii, jj, kk, QUT) = Ip[2][QUT];"),
]
tests = sys.argv[1:] or ["tests/test_oracle_pins.py", "tests/test_oracle_hydro.py", "tests/test_oracle_eos.py"]
res = []
for name, f, occ, a, b in M:
    d = os.path.join(tempfile.gettempdir(), "orc_mutant_%d" % len(res))
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    for g in os.listdir(ROOT + "/oracle"):
        if g.endswith((".c", ".h")):
            shutil.copy(ROOT + "/oracle/" + g, d)
    src = open(os.path.join(d, f)).read()
    k = -1
    for _ in range(occ + 1):
        k = src.index(a, k + 1)
    src = src[:k] + b + src[k + len(a):]
    open(os.path.join(d, f), "w").write(src)
    subprocess.run("gcc -O2 -std=c11 -fPIC -ffp-contract=off -fno-fast-math -fopenmp -w -shared -o %s/liboracle.so "
                   "%s/orc_eos.c %s/orc_hydro.c %s/orc_host.c -lm" % (d, d, d, d), shell=True, check=True)
    env = dict(os.environ, ORC_LIB=d + "/liboracle.so")
    r = subprocess.run(["python", "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider"] + tests, cwd=ROOT, env=env,
                       capture_output=True, text=True)
    last = [l for l in r.stdout.splitlines() if l.strip()][-1]
    failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
    res.append((name, r.returncode, failed[0] if failed else last))
    print("%-62s %s  %s" % (name, "caught  " if r.returncode else "SURVIVES", (failed[0] if failed else last)[:100]),
          flush=True)
bad = [n for n, rc, _ in res if rc == 0 and not n.startswith("M1 ")]
sys.exit(1 if bad else 0)
