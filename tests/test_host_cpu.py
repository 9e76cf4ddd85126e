"""Host-side pieces of the product that run without a GPU: the C ABI exports every declared
symbol, the lattice/velocity setup matches the compiled reference bit for bit, the EAM table
pipeline matches it too, and the synthetic funcfl writer agrees with its Python statement."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import REPO, load_golden, meta, product


def _declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(REPO, "include"))):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(REPO, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", src, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.append(name)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = product().lib()
    names = _declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_park_miller_known_answer():
    # Park & Miller (CACM 1988): from seed 1, the 10000th value of the minimal standard
    # generator is 1043618065.
    L = product().lib()
    s = C.c_int(1)
    for _ in range(10000):
        L.mdbMyrandom(C.byref(s))
    assert s.value == 1043618065


@pytest.mark.parametrize("pattern", ["fix", "rand"])
def test_synthetic_pattern_rejects_lists_naming_missing_atoms(pattern):
    # 10 atoms cannot hold 76 distinct neighbours (rand) or the atoms 0..75 (fix); the call
    # returns an argument error before any device work.
    P = product()
    out = np.zeros(5)
    rc = P.lib().mdbStubRun(P.STUB_PATTERNS[pattern], 10, 76, 1, 1, 0, 1, out)
    assert rc != 0 and not out.any()


@pytest.mark.parametrize("case", ["vl_lj_4", "vl_lj_5x6x7", "vl_lj_types"])
def test_lattice_and_velocities_match_reference(case):
    d = load_golden(case)
    args = [str(a) for a in d["args"]]
    kw = dict(nx=int(args[0]), ny=int(args[1]), nz=int(args[2]))
    if case == "vl_lj_types":
        kw["ntypes"] = 2
    x, v, t = product().create_lattice(**kw)
    np.testing.assert_array_equal(x, d["create_x"])
    np.testing.assert_array_equal(v, d["adjust_v"])
    np.testing.assert_array_equal(t, d["create_type"])


class EamTables(C.Structure):
    _fields_ = [("nr", C.c_int), ("nrho", C.c_int), ("nr_tot", C.c_int), ("nrho_tot", C.c_int),
                ("dr", C.c_double), ("drho", C.c_double), ("rdr", C.c_double),
                ("rdrho", C.c_double), ("cut", C.c_double), ("mass", C.c_double),
                ("rhor_spline", C.POINTER(C.c_double)), ("frho_spline", C.POINTER(C.c_double)),
                ("z2r_spline", C.POINTER(C.c_double))]


def _load_tables(path):
    L = product().lib()
    L.mdbEamLoad.argtypes = [C.c_char_p, C.POINTER(EamTables)]
    L.mdbEamFree.argtypes = [C.POINTER(EamTables)]
    t = EamTables()
    assert L.mdbEamLoad(path.encode(), C.byref(t)) == 0
    out = {k: np.ctypeslib.as_array(getattr(t, k + "_spline"),
                                    shape=(t.nrho_tot if k == "frho" else t.nr_tot,)).copy()
           for k in ("rhor", "frho", "z2r")}
    out.update(nr=t.nr, nrho=t.nr, mass=t.mass, cut=t.cut, rdr=t.rdr)
    L.mdbEamFree(C.byref(t))
    return out


def test_synthetic_funcfl_matches_python_statement(tmp_path):
    from golden.eam_tables import write_funcfl_synthetic
    a, b = str(tmp_path / "c.eam"), str(tmp_path / "py.eam")
    assert product().lib().mdbWriteFuncflSynthetic(a.encode()) == 0
    write_funcfl_synthetic(b)
    assert open(a).read() == open(b).read()


@pytest.mark.parametrize("case", ["vl_eam_synth_4", "vl_eam_cu_u3_4"])
def test_eam_splines_match_reference(case, tmp_path):
    d = load_golden(case)
    if "Cu_u3" in str(d["args"][7]):
        path = os.path.join(REPO, "tests", "golden", "Cu_u3.eam")
    else:
        path = str(tmp_path / "s.eam")
        product().lib().mdbWriteFuncflSynthetic(path.encode())
    t = _load_tables(path)
    assert t["nr"] == meta(d, "eam_nr")
    for k in ("rhor", "frho", "z2r"):
        np.testing.assert_array_equal(t[k], d["eam_%s_spline" % k])


def test_maxwell_velocities():
    """maxwell_seed: normal velocity components (Box-Muller), then adjustThermo's centre-of-mass
    removal and rescale.  The draw is reproducible, the components pass a Kolmogorov-Smirnov
    test against the standard normal and have its fourth moment (3; createAtom's uniform draw
    has 1.8), and the adjusted set has zero net momentum and exactly the requested
    temperature."""
    from scipy import stats
    P = product()
    L = P.lib()
    L.mdbMaxwellNormals.argtypes = [C.c_int, C.c_ulonglong, np.ctypeslib.ndpointer(np.float64)]
    n = 131072
    a, b, c = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
    L.mdbMaxwellNormals(n, 4242, a)
    L.mdbMaxwellNormals(n, 4242, b)
    L.mdbMaxwellNormals(n, 4243, c)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    assert stats.kstest(a, "norm").pvalue > 1e-3
    assert abs(a.mean()) < 5 / np.sqrt(a.size)
    assert abs(a.var() - 1.0) < 5 * np.sqrt(2.0 / a.size)
    assert abs(stats.kurtosis(a, fisher=False) - 3.0) < 0.05
    # through the lattice setup: LJ units, T = 1.44, 8^3 cells
    x, v, t = P.create_lattice(nx=8, ny=8, nz=8, maxwell_seed=7)
    x0, v0, _ = P.create_lattice(nx=8, ny=8, nz=8)
    np.testing.assert_array_equal(x, x0)
    v = v.reshape(-1, 3)
    assert np.abs(v.sum(0)).max() < 1e-10
    temp = (v * v).sum() / (3 * v.shape[0] - 3)
    assert abs(temp - 1.44) < 1e-12
    assert abs(stats.kurtosis(v.reshape(-1), fisher=False) - 3.0) < 0.15
    assert abs(stats.kurtosis(v0.reshape(-1), fisher=False) - 1.8) < 0.05
