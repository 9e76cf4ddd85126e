"""MD-Bench's own Verlet-list program (its main.c and every other source file unchanged) linked to
libmdgfx950 by the link-level binding of oracle/dropin/vl_gfx950.c (`make -C oracle dropin`,
oracle/_ref/mdbench-vl-gfx950): the program's thermo table against MD-Bench's own scalar CPU
build (oracle/_ref/mdbench-vl-strict) run in the same job, and against the values of
tests/sim_copper_fcc_regression.sh (step 200 of the default LJ case: T 7.961676e-01 within 1e-4
relative, P 6.721195e-01 within 1e-3).  With MDGFX950_STRICT=1 the device kernels use MD-Bench's
scalar arithmetic and the tables are identical line for line; with the default kernels the
regression tolerances hold."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(REPO, "oracle", "_ref", "mdbench-vl-gfx950")
CPU = os.path.join(REPO, "oracle", "_ref", "mdbench-vl-strict")
EAM_FILE = os.path.join(REPO, "tests", "golden", "Cu_u3.eam")
REF_T, REF_P, T_TOL, P_TOL = 7.961676e-01, 6.721195e-01, 1e-4, 1e-3
ROW = re.compile(r"^\s*(\d+)\s+([-+0-9.eE]+)\s+([-+0-9.eE]+)\s*$")

if not (os.path.exists(DROPIN) and os.path.exists(CPU)):
    pytest.skip("the MD-Bench builds under oracle/_ref are not built (make -C oracle ref dropin)",
                allow_module_level=True)


def _thermo(exe, args, strict=None):
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "16"))
    if strict is not None:
        env["MDGFX950_STRICT"] = "1" if strict else "0"
    r = subprocess.run([exe] + args, cwd=REPO, env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, (exe, r.stdout[-2000:], r.stderr[-2000:])
    rows = [m.groups() for m in map(ROW.match, r.stdout.splitlines()) if m]
    assert rows, r.stdout[-2000:]
    print("%s %s:" % (os.path.basename(exe), " ".join(args)),
          " | ".join(" ".join(t) for t in rows), flush=True)
    return rows


def _regression(rows):
    step, t, p = int(rows[-1][0]), float(rows[-1][1]), float(rows[-1][2])
    assert step == 200
    assert abs(t - REF_T) / REF_T <= T_TOL, t
    assert abs(p - REF_P) / REF_P <= P_TOL, p


def test_lj_default_case_strict_equals_cpu_build():
    cpu = _thermo(CPU, ["-n", "200"])
    gpu = _thermo(DROPIN, ["-n", "200"], strict=True)
    assert gpu == cpu
    _regression(gpu)


def test_lj_default_case_fast_kernels_meet_regression_values():
    _regression(_thermo(DROPIN, ["-n", "200"], strict=False))


def test_eam_copper_strict_equals_cpu_build():
    args = ["-f", "eam", "-e", EAM_FILE, "-n", "100", "-nx", "8", "-ny", "8", "-nz", "8"]
    assert _thermo(DROPIN, args, strict=True) == _thermo(CPU, args)
