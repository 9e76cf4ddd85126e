"""The oracle's LJ pair energy and virial, pinned by what the mathematics fixes.

MD-Bench computes LJ forces only (src/verletlist/force_lj.c:73-79 accumulates the force and
nothing else; src/common/thermo.c:83 forms a kinetic pressure).  The oracle adds the pair energy
E = sum_pairs 4 eps ((sigma/r)^12 - (sigma/r)^6) and the virial W = sum_pairs r . F_pair
(orc_vl_force_lj in oracle/oracle_vl.c, orc_cp_force_lj in oracle/oracle_cp.c, which equals the
Verlet-list sums on the same coordinates: tests/test_oracle_cp.py).  Its forces are bit-identical
to the reference's, and these tests tie energy and virial to them and to closed forms:
- minus the central-difference gradient of E equals the force (full list, half list, and a
  two-type Kob-Andersen table), ghost images moving with their atoms;
- on the perfect fcc lattice E/N equals the closed-form shell sum
  1/2 sum_s n_s 4 eps ((sigma/r_s)^12 - (sigma/r_s)^6), with the shell counts n_s found by
  enumerating lattice vectors (12, 6, 24, 12 inside r_c = 2.5 at rho* = 0.8442);
- W = -dE/ds under the isotropic scaling x -> (1+s) x of every position (W = -3V dE/dV), on a
  thermalised state and on the lattice;
- the half list gives the full list's energy and virial.
test_planted_mistakes_fail compiles copies of the oracle with one mistake each (pair weight 1
instead of 1/2, half-list local pairs weighted 1/2, 2 eps instead of 4 eps, the virial's sign)
and shows that these checks catch every one.
"""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_bind as ob
from conftest import load_golden, meta

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
H = 1e-5  # central-difference step

KA_EPS = np.array([1.0, 1.5, 1.5, 0.5])
KA_SIG = np.array([1.0, 0.8, 0.8, 0.88])


def _one_type():
    return 1, np.ones(1), np.ones(1), np.full(1, 6.25)


def _ka():
    return 2, KA_EPS, KA_SIG ** 6, (2.5 * KA_SIG) ** 2


def _thermal_state(ntypes, seed=7, sigma=0.04):
    """The 4^3-cell reference state with thermal displacements: positions (ghosts follow their
    atoms), types (80/20 for two types), grid and image shifts."""
    d = load_golden("vl_lj_4")
    nl = meta(d, "s0_nlocal")
    x0 = np.array(d["s0_x"], np.float64)
    bmap = d["s0_border_map"]
    img = x0[3 * nl:].reshape(-1, 3) - x0[:3 * nl].reshape(-1, 3)[bmap]
    rng = np.random.default_rng(seed)
    xl = x0[:3 * nl].reshape(-1, 3) + rng.normal(0.0, sigma, size=(nl, 3))
    t = (rng.random(nl) < 0.2).astype(np.int32) if ntypes == 2 else np.zeros(nl, np.int32)
    g = ob.vl_grid(4, 4, 4, meta(d, "lattice"), meta(d, "cutneigh"))
    return dict(nl=nl, xl=xl, bmap=bmap, img=img, t=t, g=g)


def _lattice_state():
    d = load_golden("vl_lj_4")
    lat = meta(d, "lattice")
    x = np.array(d["create_x"], np.float64)  # createAtom's perfect lattice
    nl = x.size // 3
    prd = (4 * lat,) * 3
    bmap, pbc = ob.setup_pbc(nl, x, prd, meta(d, "cutneigh"))
    xall = np.zeros(3 * (nl + bmap.size))
    xall[:3 * nl] = x
    xall = ob.update_pbc(nl, xall, bmap, pbc, prd)
    img = xall[3 * nl:].reshape(-1, 3) - x.reshape(-1, 3)[bmap]
    g = ob.vl_grid(4, 4, 4, lat, meta(d, "cutneigh"))
    return dict(nl=nl, xl=x.reshape(-1, 3), bmap=bmap, img=img, t=np.zeros(nl, np.int32), g=g,
                lattice=lat)


def _positions(s, xl=None):
    xl = s["xl"] if xl is None else xl
    return np.concatenate([xl, xl[s["bmap"]] + s["img"]]).reshape(-1)


def _lists(s, x, ntypes, half):
    tall = np.concatenate([s["t"], s["t"][s["bmap"]]]).astype(np.int32)
    cn = np.full(ntypes * ntypes, 2.8 ** 2)
    nn, nb = ob.build_neighbor(s["g"], s["nl"], x, tall, ntypes, cn, half)
    return tall, nn, nb


def _pair_r(s, x, nn, nb):
    """Distances of every listed pair."""
    xa = x.reshape(-1, 3)
    i = np.repeat(np.arange(s["nl"]), nn)
    return np.sqrt(((xa[i] - xa[nb]) ** 2).sum(-1)), i


def check_gradient(L, ntypes, half, seed=7):
    """-dE/dx_k by central differences equals the oracle's force on k."""
    nt, eps, sig6, cutf = _ka() if ntypes == 2 else _one_type()
    s = _thermal_state(nt, seed)
    x = _positions(s)
    tall, nn, nb = _lists(s, x, nt, half)
    f, _, _ = ob.force_lj(s["nl"], x, tall, nt, eps, sig6, cutf, nn, nb, half, L=L)
    # atoms none of whose pairs lies within 10 h of its cutoff (the energy jumps there)
    r, i = _pair_r(s, x, nn, nb)
    rc = np.sqrt(cutf[tall[i] * nt + tall[nb]])
    near = np.zeros(s["nl"], bool)
    bad = np.abs(r - rc) < 10 * H
    owner = np.where(nb < s["nl"], nb, s["bmap"][np.maximum(nb - s["nl"], 0)])
    near[i[bad]] = True
    near[owner[bad]] = True
    cand = [k for k in range(0, s["nl"], 17) if not near[k]]
    assert len(cand) >= 10
    for k in cand[:12]:
        for c in range(3):
            es = []
            for sgn in (1, -1):
                xl = s["xl"].copy()
                xl[k, c] += sgn * H
                es.append(ob.force_lj(s["nl"], _positions(s, xl), tall, nt, eps, sig6, cutf, nn,
                                      nb, half, L=L)[1])
            fd = -(es[0] - es[1]) / (2 * H)
            assert abs(fd - f[3 * k + c]) <= 1e-6 * max(1.0, abs(f[3 * k + c])), (k, c, fd,
                                                                                   f[3 * k + c])


def _fcc_shells(lat, rc):
    """Shell radii and counts of the fcc lattice inside rc, by enumerating lattice vectors
    (a/2)(i, j, k) with i + j + k even."""
    m = int(np.ceil(2 * rc / lat)) + 1
    g = np.arange(-m, m + 1)
    i, j, k = np.meshgrid(g, g, g, indexing="ij")
    v = np.stack([i, j, k], -1).reshape(-1, 3)
    v = v[(v.sum(-1) % 2 == 0) & (np.abs(v).sum(-1) > 0)]
    n2 = (v ** 2).sum(-1)  # |r|^2 = n2 (a/2)^2
    keep = n2 * (lat / 2) ** 2 < rc * rc
    vals, counts = np.unique(n2[keep], return_counts=True)
    return np.sqrt(vals) * lat / 2, counts


def check_lattice_shell_sum(L, half):
    s = _lattice_state()
    lat = s["lattice"]
    radii, counts = _fcc_shells(lat, 2.5)
    assert counts.tolist() == [12, 6, 24, 12]
    sr6 = radii ** -6.0
    e_atom = 0.5 * (counts * 4.0 * sr6 * (sr6 - 1.0)).sum()
    # W/N = 1/2 sum_s n_s r_s F(r_s), F(r) = 48 (r^-12 - r^-6/2) / r along r: r.F = 48 sr6 (sr6 - 1/2)
    w_atom = 0.5 * (counts * 48.0 * sr6 * (sr6 - 0.5)).sum()
    x = _positions(s)
    nt, eps, sig6, cutf = _one_type()
    tall, nn, nb = _lists(s, x, nt, half)
    f, e, w = ob.force_lj(s["nl"], x, tall, nt, eps, sig6, cutf, nn, nb, half, L=L)
    n = s["nl"]
    assert abs(e / n - e_atom) <= 1e-12 * abs(e_atom), (e / n, e_atom)
    assert abs(w / n - w_atom) <= 1e-12 * abs(w_atom), (w / n, w_atom)
    assert np.abs(f[:3 * n]).max() < 1e-12  # fcc symmetry


def check_virial_scaling(L, half, ntypes=1, seed=7):
    """W = -dE/ds for x -> (1+s) x (every position, ghosts included: the box scales too)."""
    nt, eps, sig6, cutf = _ka() if ntypes == 2 else _one_type()
    s = _thermal_state(nt, seed)
    x = _positions(s)
    tall, nn, nb = _lists(s, x, nt, half)
    _, e0, w = ob.force_lj(s["nl"], x, tall, nt, eps, sig6, cutf, nn, nb, half, L=L)
    r, i = _pair_r(s, x, nn, nb)
    rc = np.sqrt(cutf[tall[i] * nt + tall[nb]])
    hs = 1e-7
    assert np.abs(r - rc).min() > 10 * hs * rc.max()  # no pair crosses its cutoff
    es = [ob.force_lj(s["nl"], x * (1 + sgn * hs), tall, nt, eps, sig6, cutf, nn, nb, half,
                      L=L)[1] for sgn in (1, -1)]
    dEds = (es[0] - es[1]) / (2 * hs)
    assert abs(w + dEds) <= 1e-6 * abs(w), (w, -dEds)


def check_half_equals_full(L, ntypes=1):
    nt, eps, sig6, cutf = _ka() if ntypes == 2 else _one_type()
    s = _thermal_state(nt)
    x = _positions(s)
    out = []
    for half in (0, 1):
        tall, nn, nb = _lists(s, x, nt, half)
        f, e, w = ob.force_lj(s["nl"], x, tall, nt, eps, sig6, cutf, nn, nb, half, L=L)
        out.append((f[:3 * s["nl"]], e, w))
    (f0, e0, w0), (f1, e1, w1) = out
    assert abs(e1 - e0) <= 1e-12 * abs(e0), (e0, e1)
    assert abs(w1 - w0) <= 1e-12 * abs(w0), (w0, w1)


@pytest.mark.parametrize("ntypes", [1, 2])
@pytest.mark.parametrize("half", [0, 1])
def test_energy_gradient_is_the_force(ntypes, half):
    check_gradient(None, ntypes, half)


@pytest.mark.parametrize("half", [0, 1])
def test_lattice_energy_and_virial_equal_shell_sums(half):
    check_lattice_shell_sum(None, half)


@pytest.mark.parametrize("ntypes", [1, 2])
@pytest.mark.parametrize("half", [0, 1])
def test_virial_is_minus_volume_derivative(ntypes, half):
    check_virial_scaling(None, half, ntypes)


@pytest.mark.parametrize("ntypes", [1, 2])
def test_half_list_energy_equals_full_list(ntypes):
    check_half_equals_full(None, ntypes)


# one mistake each in a copy of oracle_vl.c, and the checks that must catch it
_WGT = "double wgt = (half && j < nlocal) ? 1.0 : 0.5;"
MUTANTS = {
    "full-list pair weight 1": (_WGT, "double wgt = (half && j < nlocal) ? 1.0 : 1.0;",
                                ["gradient-full", "lattice-full", "half-full"]),
    "half-list local pair weight 1/2": (_WGT, "double wgt = (half && j < nlocal) ? 0.5 : 0.5;",
                                        ["gradient-half", "lattice-half", "half-full"]),
    "2 eps instead of 4 eps": ("e += wgt * 4.0 * epsilon[tij]", "e += wgt * 2.0 * epsilon[tij]",
                               ["gradient-full", "gradient-half", "lattice-full", "virial-full"]),
    "virial sign": ("w += wgt * rsq * force;", "w -= wgt * rsq * force;",
                    ["lattice-full", "virial-full", "virial-half"]),
}

CHECKS = {
    "gradient-full": lambda L: check_gradient(L, 1, 0),
    "gradient-half": lambda L: check_gradient(L, 1, 1),
    "lattice-full": lambda L: check_lattice_shell_sum(L, 0),
    "lattice-half": lambda L: check_lattice_shell_sum(L, 1),
    "virial-full": lambda L: check_virial_scaling(L, 0),
    "virial-half": lambda L: check_virial_scaling(L, 1),
    "half-full": lambda L: check_half_equals_full(L),
}


def _build_mutant(tmp, old, new):
    src = open(os.path.join(ORACLE_DIR, "oracle_vl.c")).read()
    assert src.count(old) == 1, old
    mut = os.path.join(tmp, "oracle_vl.c")
    with open(mut, "w") as fh:
        fh.write(src.replace(old, new))
    others = [os.path.join(ORACLE_DIR, f) for f in sorted(os.listdir(ORACLE_DIR))
              if f.endswith(".c") and f != "oracle_vl.c"]
    so = os.path.join(tmp, "liboracle_mut.so")
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                    "-D_GNU_SOURCE", "-I" + ORACLE_DIR, "-shared", "-o", so, mut] + others +
                   ["-lm"], check=True)
    return ob.load(so)


@pytest.mark.parametrize("name", list(MUTANTS))
def test_planted_mistakes_fail(name):
    old, new, must_fail = MUTANTS[name]
    with tempfile.TemporaryDirectory() as tmp:
        L = _build_mutant(tmp, old, new)
        for check in must_fail:
            with pytest.raises(AssertionError):
                CHECKS[check](L)
