"""The oracle's pins catch plausible mistakes (CPU only).

Each mutant is a copy of oracle/sph_oracle.cpp with one plausible slip planted
in it (a density taken from the wrong particle, a sign, a dropped factor, a
wrong branch bound, the textbook leap-frog in place of the reference's Q12
update, ...), built with the oracle's own flags into a temporary directory.
The CPU pin tests then run against that build (tests/test_oracle.py and
tests/test_oracle_pins.py, through SPH_ORACLE_LIB), and at least one of them
must fail. The source under oracle/ is never modified.
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "sph_oracle.cpp")
sys.path.insert(0, ROOT)
from oracle import build as oracle_build  # noqa: E402

# (name, original text, planted text, occurrence index among equal matches)
MUTANTS = [
    ("normals_rho_i", "nn = add(nn, vdiv(mul(r, q), O.density[j]));", "nn = add(nn, vdiv(mul(r, q), O.density[i]));", 0),
    ("normals_no_h", "O.normal[i] = mul(o, H);", "O.normal[i] = o;", 0),
    ("visc_rho_i", "(H - l) / dj));", "(H - l) / di));", 0),
    ("curvature_sign", "fcurv = add(fcurv, mul(sub(ni, O.normal[j]), corr));",
     "fcurv = add(fcurv, mul(sub(O.normal[j], ni), corr));", 0),
    ("spline_branch", "if (l <= T(0.5f) * H) {", "if (l <= T(0.25f) * H) {", 0),
    ("spline_term_sign", "stk = T(2) * t2 - T(pc.st_term);", "stk = T(2) * t2 + T(pc.st_term);", 0),
    ("cohesion_no_mass", "fcoh = mul(fcoh, M * T(pc.st_main));", "fcoh = mul(fcoh, T(pc.st_main));", 0),
    ("pred_density_current_pos", "V3<T> r = sub(xi, O.pos_pred[nl[(size_t)k * n + p]]);",
     "V3<T> r = sub(xi, v3<T>(pos_s[nl[(size_t)k * n + p]]));", 0),
    ("pressure_not_accumulated", "O.pressure[p] += dv * delta;", "O.pressure[p] = dv * delta;", 0),
    ("pred_density_boundary_mass", "pdb *= wde;", "pdb *= wde * M;", 0),
    ("boundary_pressure_no_2", "fb = add(fb, smul(T(b_phi_s[j]) * (T(2) * C), gw));",
     "fb = add(fb, smul(T(b_phi_s[j]) * C, gw));", 0),
    ("fluid_force_no_2", "V3<T> t = mul(mul(mul(mul(r, sg / rn), sqr(H - rn)), T(2)), C);",
     "V3<T> t = mul(mul(mul(r, sg / rn), sqr(H - rn)), C);", 0),
    ("fluid_force_no_psi", "f = mul(f, mass * T(b_phi[bi]));", "f = mul(f, mass);", 0),
    ("fluid_force_sign", "V3<T> r = sub(v3<T>(pos_s[j]), pi);", "V3<T> r = sub(pi, v3<T>(pos_s[j]));", 0),
    ("wc_boundary_no_2", "t = mul(t, T(2));", "t = mul(t, T(1));", 0),
    ("wc_pressure_rho_i", "C + prj / sqr(dj));", "C + prj / sqr(di));", 0),
    ("wc_visc_rho_i", "h - rn), dj));", "h - rn), di));", 0),
    ("wc_cohesion_no_c1", "fcoh = mul(fcoh, T(wc.st_main));", "fcoh = mul(fcoh, T(1));", 0),
    ("wc_normals_rho_i", "nn = add(nn, vdiv(mul(r, q), O.density[j]));", "nn = add(nn, vdiv(mul(r, q), O.density[i]));", 1),
    ("wc_leapfrog_textbook", "x = add(x, smul(dt, cl_clamp3(add(vh, smul(dt, a)), -maxv, maxv)));",
     "x = add(x, smul(dt, cl_clamp3(vh, -maxv, maxv)));", 0),
    ("wc_vt_full_step", "V3<T> vt = add(vh, mul(smul(dt, a), T(0.5f)));", "V3<T> vt = add(vh, smul(dt, a));", 0),
    ("wc_no_restitution", "T rest = (T(1) * dd) / (dt * vn);", "T rest = T(0) * dd;", 0),
    ("wc_tait_exponent", "O.pressure[i] = cl_max(T(0), k_gas * (pown7(d / rho0) - T(1)));",
     "O.pressure[i] = cl_max(T(0), k_gas * (cube(d / rho0) - T(1)));", 0),
]


def plant(src, old, new, which):
    at = -1
    for _ in range(which + 1):
        at = src.index(old, at + 1)
    return src[:at] + new + src[at + len(old):]


def build_and_check(tmp, name, old, new, which):
    with open(SRC) as f:
        src = f.read()
    mutated = plant(src, old, new, which)
    assert mutated != src
    cpp = os.path.join(tmp, name + ".cpp")
    lib = os.path.join(tmp, name + ".so")
    with open(cpp, "w") as f:
        f.write(mutated)
    subprocess.run(["g++", *oracle_build.FLAGS, cpp, "-o", lib], check=True, capture_output=True)
    env = dict(os.environ, SPH_ORACLE_LIB=lib, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py"),
                        os.path.join(ROOT, "tests", "test_oracle.py"), "-k", "not full_size"],
                       env=env, capture_output=True, text=True, cwd=ROOT)
    failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
    return name, r.returncode, failed


def test_every_mutant_fails_a_pin():
    with open(SRC) as f:
        src = f.read()
    for name, old, new, which in MUTANTS:
        assert src.count(old) > which, name
    with tempfile.TemporaryDirectory() as tmp, cf.ThreadPoolExecutor(max_workers=6) as ex:
        results = list(ex.map(lambda m: build_and_check(tmp, *m), MUTANTS))
    survivors = [name for name, rc, _ in results if rc == 0]
    for name, rc, failed in results:
        print(f"{name:30s} rc={rc} {failed[:1]}")
    assert not survivors, survivors
