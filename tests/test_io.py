"""MD-Bench's dump format (src/verletlist/atom.c:529-656): the writer and reader of the host
library round-trip exactly, and a state written here starts MD-Bench itself (its strict build in
oracle/_ref, run with -i) and the oracle on the same trajectory."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_bind as ob
from conftest import REPO, product

REF_VL = os.path.join(REPO, "oracle", "_ref", "mdbench-vl-strict")
LAT = (4 / 0.8442) ** (1 / 3)


def _state(nx=4, seed=1, sigma=0.05):
    P = product()
    x, v, t = P.create_lattice(nx=nx, ny=nx, nz=nx)
    L = nx * LAT
    rng = np.random.default_rng(seed)
    x = np.mod(x + rng.normal(0, sigma, x.size), L)
    return x, v, t, [0.0, L, 0.0, L, 0.0, L]


def test_dump_round_trip_is_exact(tmp_path):
    P = product()
    x, v, t, box = _state()
    path = str(tmp_path / "s.dmp")
    P.write_dmp(path, x, v, t, box, timestep=7)
    x2, v2, t2, box2, nt = P.read_dmp(path, x.size // 3)
    np.testing.assert_array_equal(x2, x)
    np.testing.assert_array_equal(v2, v)
    np.testing.assert_array_equal(t2, t)
    np.testing.assert_array_equal(box2, box)
    assert nt == t.max() + 1
    head = open(path).read().splitlines()[:9]
    assert head[0] == "ITEM: TIMESTEP" and head[1] == "7"
    assert head[2] == "ITEM: NUMBER OF ATOMS" and head[4] == "ITEM: BOX BOUNDS pp pp pp"
    assert head[8] == "ITEM: ATOMS id type x y z vx vy vz"


def test_reader_rejects_malformed_files(tmp_path):
    P = product()
    bad = tmp_path / "bad.dmp"
    bad.write_text("ITEM: TIMESTEP\n0\nITEM: SOMETHING ELSE\n")
    x, v, t = np.zeros(30), np.zeros(30), np.zeros(10, np.int32)
    import ctypes as C
    rc = P.lib().mdbReadDmp(str(bad).encode(), 10, x, v, t, np.zeros(6), C.byref(C.c_int()))
    assert rc < 0
    head = "ITEM: TIMESTEP\n0\nITEM: NUMBER OF ATOMS\n3\nITEM: BOX BOUNDS pp pp pp\n" \
           "0 4\n0 4\n0 4\nITEM: ATOMS id type x y z vx vy vz\n"
    good = ["1 0 1 1 1 0 0 0", "2 1 2 2 2 0 0 0", "3 0 3 3 3 0 0 0"]
    cases = {
        "ok": good,
        "negative type": ["1 0 1 1 1 0 0 0", "2 -1 2 2 2 0 0 0", "3 0 3 3 3 0 0 0"],
        "type beyond the tables": ["1 0 1 1 1 0 0 0", "2 8 2 2 2 0 0 0", "3 0 3 3 3 0 0 0"],
        "repeated id": ["1 0 1 1 1 0 0 0", "1 1 2 2 2 0 0 0", "3 0 3 3 3 0 0 0"],
        "id beyond the count": ["1 0 1 1 1 0 0 0", "4 1 2 2 2 0 0 0", "3 0 3 3 3 0 0 0"],
        "missing atom line": good[:2],
    }
    for name, lines in cases.items():
        bad.write_text(head + "\n".join(lines) + "\n")
        nt = C.c_int()
        rc = P.lib().mdbReadDmp(str(bad).encode(), 10, x, v, t, np.zeros(6), C.byref(nt))
        if name == "ok":
            assert rc == 3 and nt.value == 2, rc
            assert t[:3].tolist() == [0, 1, 0]
        else:
            assert rc < 0, name


def _reference_thermo(path, nsteps):
    out = subprocess.run([REF_VL, "-i", path, "-n", str(nsteps)], capture_output=True, text=True,
                         timeout=120, cwd=os.path.dirname(path), errors="replace")
    assert out.returncode == 0, out.stdout[-500:]
    rows = re.findall(r"^\s+(\d+)\s+([-+0-9.eE]+)\s+([-+0-9.eE]+)\s*$", out.stdout, flags=re.M)
    return {int(s): (float(t), float(p)) for s, t, p in rows}


@pytest.mark.skipif(not os.path.exists(REF_VL), reason="MD-Bench build (oracle/_ref) absent")
def test_mdbench_reads_the_dump_and_matches_the_oracle(tmp_path):
    x, v, t, box = _state()
    path = str(tmp_path / "s.dmp")
    product().write_dmp(path, x, v, t, box)
    th = _reference_thermo(path, 40)
    nt = int(t.max()) + 1
    one = np.ones(nt * nt)
    _, _, _, ks = ob.vl_run_box(np.array(box), x, v, t, nt, one, one, one * 6.25, one * 2.8 ** 2,
                                2.8, 0, 40, 20, 0.005, 0.0025, 1.0)
    n = x.size // 3
    T = ks / (3 * n - 3)
    assert th[0][0] == float("%.6e" % T[0])
    assert th[40][0] == float("%.6e" % T[40])
