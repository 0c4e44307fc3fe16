"""The closed-form pins catch plausible misreadings of the reference.

Each mutant is a copy of oracle/tc_oracle.c with one deliberate mistake of
the kind a restatement could make (a swapped constant, a wrong index, a
dropped factor, a flipped sign).  It is compiled on the CPU into a temporary
library, and tests/test_oracle_pins.py and tests/test_oracle_bcs.py are run
against it (ORACLE_LIB): every mutant must make at least one pin fail, while
the unmodified oracle passes them all (those two files themselves).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "tc_oracle.c")

# (name, reference routine, text in tc_oracle.c, its mutation)
#
# Not a mutant: swapping la and lb of smoother_vis's first pass
# (smoother.f90:180-183).  (lb - la) cos(theta_j) only changes sign, and the
# set {cos((2j-1) pi / 2N)} is symmetric, so the swap reverses the order of
# the pass's step sizes and leaves their product unchanged.  That pass always
# runs to its end (Nvis_store = 0), so the order shows only in rounding; the
# order is pinned where a pass can stop early (the ntemp mutants below).
MUTANTS = [
    ("smoother_vis ntemp counted from nn", "smoother.f90:237",
     "int ntemp = Niter - nn + 1;", "int ntemp = nn;"),
    ("smoother_vis without the 0.99", "smoother.f90:245",
     "double dts = dtv[cc] * 0.99;", "double dts = dtv[cc];"),
    ("res_ref normalised by (n-1)^2", "iterative_solver.f90:286",
     "acc / (double)(nxl - 2) / (double)(nzl - 2);", "acc / (double)(nxl - 1) / (double)(nzl - 1);"),
    ("iterate_stream ntemp counted from n", "iterative_solver.f90:873",
     "double ntemp = (double)(Niter - n + 1);", "double ntemp = (double)n;"),
    ("iterate_stream lb from ratio, not ratio^2", "iterative_solver.f90:883",
     "else lb = -2.0 * (r * r) / (1.0 + r * r);", "else lb = -2.0 * r / (1.0 + r);"),
    ("visf depth factor from z instead of 1-z", "momentum.f90:8",
     "return exp(-s->Tvis[c] * log(s->p.visT) + (1.0 - zgk) * log(s->p.visP));",
     "return exp(-s->Tvis[c] * log(s->p.visT) + zgk * log(s->p.visP));"),
    ("eigI Trss scaled by dr^2 ds", "iterative_solver.f90:684",
     "s->Trss_max = s->p.Trrs_base / s->drds2g[level];", "s->Trss_max = s->p.Trrs_base / s->dr2dsg[level];"),
    ("test_eigenvalues without the ratio reset", "iterative_solver.f90:687",
     "for (size_t c = 0; c < s->plane; ++c) s->ratio[L][c] = 1e99;", "(void)L;"),
    ("compute_RHS with +RaC", "iterative_solver.f90:583",
     "rhs = rhs - s->p.RaC[j - 1] * Cx;", "rhs = rhs + s->p.RaC[j - 1] * Cx;"),
    ("strain without the factor 2 on the normal strains", "momentum.f90:50",
     "st[c] = sqrt(0.5 * (2.0 * (ux * ux + wz * wz) + a * a));", "st[c] = sqrt(0.5 * ((ux * ux + wz * wz) + a * a));"),
    ("energy_space with -grad K . grad T", "energy.f90:79",
     "s->conduct[c] * (Txx + Tzz) + (kx * Tx + kz * Tz)", "s->conduct[c] * (Txx + Tzz) - (kx * Tx + kz * Tz)"),
    ("stability_RK4 with the RK2 diffusive bound", "energy.f90:627-628",
     "else { dtA = 2.8 / eA_max; dtD = 2.8 / eD_max; }", "else { dtA = 2.8 / eA_max; dtD = 2.0 / eD_max; }"),
    ("smoother_time first pass for [1,4]", "smoother.f90:296-297",
     "const double lb = -5.0, la = -1.0;", "const double lb = -4.0, la = -1.0;"),
    ("smoother_time without courantC", "smoother.f90:323",
     "double dts = s->dts[cc] * s->p.courantC;", "double dts = s->dts[cc];"),
    ("enforceBCs top ghosts copied, not reflected about 0", "energy.f90:470",
     "T0[IX(s, i, nz + k)] = -T0[IX(s, i, nz - k)];", "T0[IX(s, i, nz + k)] = T0[IX(s, i, nz - k)];"),
    ("enforceBCs_comp bottom ghosts from the row one further in", "composition.f90:800",
     "C0[IX(s, ii, -i)] = C0[IX(s, ii, i)];", "C0[IX(s, ii, -i)] = C0[IX(s, ii, i + 1)];"),
]


def _build(tmp, idx, old, new):
    src = open(SRC).read()
    assert src.count(old) == 1, f"mutation site not unique or gone: {old!r}"
    d = os.path.join(tmp, f"m{idx}")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, "tc_oracle.c")
    with open(path, "w") as f:
        f.write(src.replace(old, new))
    lib = os.path.join(d, "liboracle.so")
    cmd = ["gcc", "-O1", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-I", os.path.join(ROOT, "oracle"),
           "-o", lib, path, "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return lib


def _run_pins(lib):
    env = dict(os.environ, ORACLE_LIB=lib, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py"),
                        os.path.join(ROOT, "tests", "test_oracle_bcs.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-400:]


def test_every_mutant_fails_a_pin(tmp_path):
    def one(idx):
        name, _, old, new = MUTANTS[idx]
        return name, _run_pins(_build(str(tmp_path), idx, old, new))

    survivors = []
    with cf.ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        for name, (rc, tail) in ex.map(one, range(len(MUTANTS))):
            print(f"{name}: pins rc={rc}", flush=True)
            if rc == 0:
                survivors.append(name)
            else:
                assert "failed" in tail, tail  # a pin failed (not a collection or build error)
    assert not survivors, survivors


@pytest.mark.parametrize("name,ref,old,new", MUTANTS)
def test_mutation_site_exists(name, ref, old, new):
    assert open(SRC).read().count(old) == 1
