"""The CPU oracle (Verlet-list path) against the compiled reference's golden vectors.

Golden vectors come from MD-Bench's own Verlet-list sources built with strict IEEE flags
(tests/golden/make_golden.py).  Bins, ghost maps and neighbour lists are integer work and must be
identical; forces and trajectories of the strict restatement must be bit-identical too.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_bind as ob
from conftest import load_golden, meta

LJ_CASES = ["vl_lj_4", "vl_lj_4_half", "vl_lj_5x6x7", "vl_lj_types"]
EAM_CASES = ["vl_eam_synth_4", "vl_eam_cu_u3_4"]


def _dims(d):
    args = [str(a) for a in d["args"]]
    return int(args[0]), int(args[1]), int(args[2]), int(args[3])


def _grid(d):
    nx, ny, nz, _ = _dims(d)
    return ob.vl_grid(nx, ny, nz, meta(d, "lattice"), meta(d, "cutneigh"))


@pytest.mark.parametrize("case", LJ_CASES + EAM_CASES)
def test_grid_matches_reference(case):
    d = load_golden(case)
    g = _grid(d)
    for k in ["mbinx", "mbiny", "mbinz", "pad_x", "pad_y", "pad_z", "nstencil"]:
        assert getattr(g, k) == meta(d, k), k
    for k in ["binsizex", "binsizey", "binsizez", "bininvx", "bininvy", "bininvz"]:
        assert getattr(g, k) == meta(d, k), k
    assert list(g.stencil[:g.nstencil]) == list(d["stencil"])


@pytest.mark.parametrize("case", LJ_CASES + EAM_CASES)
def test_ghost_map_bins_and_lists(case):
    d = load_golden(case)
    g = _grid(d)
    nlocal = meta(d, "s0_nlocal")
    x0 = d["s0_x"][:3 * nlocal]
    bmap, pbc = ob.setup_pbc(nlocal, x0, (g.xprd, g.yprd, g.zprd), g.cutneigh)
    assert bmap.size == meta(d, "s0_nghost")
    np.testing.assert_array_equal(bmap, d["s0_border_map"])
    np.testing.assert_array_equal(pbc, d["s0_pbc"])
    x = np.zeros(3 * (nlocal + bmap.size))
    x[:3 * nlocal] = x0
    x = ob.update_pbc(nlocal, x, bmap, pbc, (g.xprd, g.yprd, g.zprd))
    np.testing.assert_array_equal(x, d["s0_x"])  # ghost positions bit-identical
    start, atoms = ob.binatoms(g, x)
    np.testing.assert_array_equal(np.diff(start), d["s0_bincount"])
    np.testing.assert_array_equal(atoms, d["s0_bins"])
    half = _dims(d)[3]
    numneigh, nb = ob.build_neighbor(g, nlocal, x, d["s0_type"], meta(d, "ntypes"),
                                     d["cutneighsq"], half)
    np.testing.assert_array_equal(numneigh, d["s0_numneigh"])
    np.testing.assert_array_equal(nb, d["s0_neighbors"])  # same order, not only same sets


@pytest.mark.parametrize("case", LJ_CASES)
def test_lj_force_bit_identical(case):
    d = load_golden(case)
    nlocal = meta(d, "s0_nlocal")
    half = _dims(d)[3]
    f, e, w = ob.force_lj(nlocal, d["s0_x"], d["s0_type"], meta(d, "ntypes"), d["epsilon"],
                          d["sigma6"], d["cutforcesq"], d["s0_numneigh"], d["s0_neighbors"], half)
    np.testing.assert_array_equal(f[:3 * nlocal], d["s0_f"])


@pytest.mark.parametrize("case", LJ_CASES + EAM_CASES)
def test_trajectory_matches_reference(case):
    d = load_golden(case)
    nx, ny, nz, half = _dims(d)
    args = [str(a) for a in d["args"]]
    ntimes, reneigh = int(args[4]), int(args[5])
    tab = None
    if args[6] == "eam":
        tab = ob.EamTables(_eam_path(args[7]))
    x, v, f, ks = ob.vl_run(nx, ny, nz, meta(d, "lattice"), d["create_x"], d["adjust_v"],
                            d["create_type"], meta(d, "ntypes"), d["epsilon"], d["sigma6"],
                            d["cutforcesq"], d["cutneighsq"], meta(d, "cutneigh"), half, ntimes,
                            reneigh, meta(d, "dt"), meta(d, "dtforce"), meta(d, "mass"), eam=tab)
    nlocal = meta(d, "end_nlocal")
    np.testing.assert_array_equal(ks, d["traj_ksum"])
    np.testing.assert_array_equal(x, d["end_x"][:3 * nlocal])
    np.testing.assert_array_equal(v, d["end_v"])
    np.testing.assert_array_equal(f, d["end_f"])


def _eam_path(arg):
    """EAM tables: the synthetic table is regenerated here; Cu_u3 comes from its fixture copy."""
    import tempfile
    from golden.eam_tables import write_funcfl_synthetic
    if "Cu_u3" in arg:
        return os.path.join(os.path.dirname(__file__), "golden", "Cu_u3.eam")
    path = os.path.join(tempfile.gettempdir(), "mdb_cu_synth_%d.eam" % os.getpid())
    write_funcfl_synthetic(path)
    return path


@pytest.mark.parametrize("case", EAM_CASES)
def test_eam_tables_and_force(case):
    d = load_golden(case)
    args = [str(a) for a in d["args"]]
    tab = ob.EamTables(_eam_path(args[7]))
    assert tab.eam.nr == meta(d, "eam_nr") and tab.eam.nrho == meta(d, "eam_nrho")
    assert tab.eam.nr_tot == meta(d, "eam_nr_tot")
    for w in ["frho", "rhor", "z2r"]:
        np.testing.assert_array_equal(tab.spline(w), d["eam_%s_spline" % w])
    nlocal = meta(d, "s0_nlocal")
    cutf = d["cutforcesq"][0]
    f, fp, e = ob.force_eam(tab, nlocal, d["s0_x"], d["s0_border_map"], cutf, d["s0_numneigh"],
                            d["s0_neighbors"])
    np.testing.assert_array_equal(f, d["s0_f"])


@pytest.mark.parametrize("case", EAM_CASES)
def test_eam_atom_energy_sums_to_total_and_is_the_force_potential(case):
    """Per-atom EAM energies F(rho_i) + 1/2 sum_j phi(r_ij): they add up to the total energy, and
    central differences of their sum reproduce the forces (moving an atom moves its ghost images
    with it; the neighbour list stays valid for the small step)."""
    d = load_golden(case)
    args = [str(a) for a in d["args"]]
    tab = ob.EamTables(_eam_path(args[7]))
    nlocal = meta(d, "s0_nlocal")
    cutf = d["cutforcesq"][0]
    x0 = np.array(d["s0_x"], np.float64)
    bmap = d["s0_border_map"]
    nn, nb = d["s0_numneigh"], d["s0_neighbors"]
    f, ea = ob.force_eam_atom(tab, nlocal, x0, bmap, cutf, nn, nb)
    f2, _, e = ob.force_eam(tab, nlocal, x0, bmap, cutf, nn, nb)
    np.testing.assert_array_equal(f, f2)
    assert abs(ea.sum() - e) <= 1e-12 * abs(e)
    # thermal displacements so that forces are not zero by symmetry
    rng = np.random.default_rng(5)
    x = x0.copy()
    dx = rng.normal(0.0, 0.05, size=(nlocal, 3))
    x[:3 * nlocal] += dx.reshape(-1)
    img = x0[3 * nlocal:].reshape(-1, 3) - x0[:3 * nlocal].reshape(-1, 3)[bmap]
    x[3 * nlocal:] = (x[:3 * nlocal].reshape(-1, 3)[bmap] + img).reshape(-1)
    f, _ = ob.force_eam_atom(tab, nlocal, x, bmap, cutf, nn, nb)
    h = 1e-5
    for k in [0, 7, nlocal // 2, nlocal - 1]:
        for c in range(3):
            es = []
            for sgn in (1, -1):
                xs = x.copy()
                xs[3 * k + c] += sgn * h
                ghosts = np.nonzero(bmap == k)[0]
                xs[3 * (nlocal + ghosts) + c] += sgn * h
                es.append(ob.force_eam_atom(tab, nlocal, xs, bmap, cutf, nn, nb)[1].sum())
            fd = -(es[0] - es[1]) / (2 * h)
            assert abs(fd - f[3 * k + c]) <= 1e-6 * max(1.0, abs(f[3 * k + c])), (k, c, fd)


def _bins(g, x):
    L = ob.lib()
    return np.array([L.orc_vl_coord2bin(C.byref(g), *x[3 * i:3 * i + 3])
                     for i in range(x.size // 3)])


def test_sort_atoms_is_the_stable_bin_order():
    """sortAtom (neighbor.c:466-536): the new order is the one permutation that makes the bins
    non-decreasing and keeps index order inside a bin, and v, f and type travel with x."""
    d = load_golden("vl_lj_4")
    g = _grid(d)
    rng = np.random.default_rng(11)
    n = 400
    x = rng.uniform(0.0, g.xprd, 3 * n)
    v, f = rng.normal(size=3 * n), rng.normal(size=3 * n)
    t = np.arange(n, dtype=np.int32)  # types as atom ids
    xs, vs, fs, ts = ob.vl_sort_atoms(g, x, v, f, t)
    perm = ts.astype(np.int64)
    assert sorted(perm.tolist()) == list(range(n))
    for a, b in ((xs, x), (vs, v), (fs, f)):
        np.testing.assert_array_equal(a.reshape(-1, 3), b.reshape(-1, 3)[perm])
    b = _bins(g, x)[perm]
    assert np.all(np.diff(b) >= 0)
    same = np.diff(b) == 0
    assert np.all(np.diff(perm)[same] > 0)
    # the binning behind it is the golden-pinned one: bin-major order of orc_vl_binatoms
    start, atoms = ob.binatoms(g, x)
    np.testing.assert_array_equal(perm, atoms)
    # sorting again changes nothing
    again = ob.vl_sort_atoms(g, xs, vs, fs, ts)
    for a, b2 in zip(again, (xs, vs, fs, ts)):
        np.testing.assert_array_equal(a, b2)


def test_resorted_run_follows_the_unsorted_one():
    """A run that re-sorts (at setup and every 7 steps, rebuilding every 5) is the unsorted run
    with its atoms relabelled: with one parameter set for every type pair and the types as atom
    ids, each atom ends where the unsorted run puts it (to the rounding that the changed order of
    the force sums makes) and the kinetic-energy trace agrees."""
    d = load_golden("vl_lj_4")
    nx, ny, nz, half = _dims(d)
    n = d["create_x"].size // 3
    nt = n
    one = np.ones(nt * nt)
    ids = np.arange(n, dtype=np.int32)
    common = (nx, ny, nz, meta(d, "lattice"), d["create_x"], d["adjust_v"], ids, nt,
              one * d["epsilon"][0], one * d["sigma6"][0], one * d["cutforcesq"][0],
              one * d["cutneighsq"][0], meta(d, "cutneigh"), 0, 30, 5, meta(d, "dt"),
              meta(d, "dtforce"), meta(d, "mass"))
    x0, v0, f0, k0 = ob.vl_run(*common)
    x1, v1, f1, k1, t1 = ob.vl_run_resort(*common, 7)
    assert sorted(t1.tolist()) == list(range(n)) and not np.array_equal(t1, ids)
    np.testing.assert_allclose(k1, k0, rtol=1e-12)
    np.testing.assert_allclose(x1.reshape(-1, 3), x0.reshape(-1, 3)[t1], rtol=0, atol=1e-10)
    np.testing.assert_allclose(v1.reshape(-1, 3), v0.reshape(-1, 3)[t1], rtol=0, atol=1e-10)
    np.testing.assert_allclose(f1.reshape(-1, 3), f0.reshape(-1, 3)[t1], rtol=0, atol=1e-8)
