"""The cluster-pair oracle against MD-Bench's own cluster-pair builds (golden vectors).

1x2 (the plain-C reference kernel build), 4x4 (AVX2 4xN build) and 4x8 (AVX-512 4xN build) are
compared on: cluster membership and bounding boxes, cluster bins, ghost clusters, neighbour list
contents and order (masked entries first), and -- for 1x2, whose force kernel is FP64 exact --
forces and whole trajectories bit for bit.  The 4x4/4x8 force of the oracle is pinned by the
Verlet-list oracle on the same coordinates: both sum the same pair set.
"""
import numpy as np
import pytest

import oracle_bind as ob
from conftest import load_golden, meta

SHAPES = {"m1n2": (1, 2), "m4n4": (4, 4), "m4n8": (4, 8)}
CASES = [(v, s) for v in SHAPES for s in ("4", "4_half", "5x6x7")]


def _setup(d, variant):
    args = [int(str(a)) for a in d["args"]]
    nx, ny, nz, half = args[:4]
    lat = meta(d, "lattice")
    M, N = SHAPES[variant]
    cp = ob.CP(M, N, (nx * lat, ny * lat, nz * lat), meta(d, "cutneigh"), d["atoms_x"],
               d["atoms_v"], d["atoms_type"])
    return cp, half, args


def _nslots(d, tag="s0"):
    return (meta(d, tag + "_ncj") + meta(d, tag + "_nghost_clusters") + 1) * meta(d, "cluster_n")


@pytest.mark.parametrize("variant,size", CASES)
def test_clusters_ghosts_lists_match_reference(variant, size):
    name = "cp_%s_%s" % (variant, size)
    d = load_golden(name)
    cp, half, _ = _setup(d, variant)
    cp.reneighbour(half)
    inf = cp.info()
    assert inf["nci"] == meta(d, "s0_nci") and inf["ncj"] == meta(d, "s0_ncj")
    assert inf["nghost"] == meta(d, "s0_nghost_clusters")
    assert inf["dummy_cj"] == meta(d, "s0_dummy_cj")
    np.testing.assert_array_equal(cp.get("ci_natoms"), d["s0_ci_natoms"])
    np.testing.assert_array_equal(cp.get("ci_bin"), d["s0_ci_bin"])
    np.testing.assert_array_equal(cp.get("ci_bbox"), d["s0_ci_bbox"])
    np.testing.assert_array_equal(cp.get("cj_natoms"), d["s0_cj_natoms"])
    np.testing.assert_array_equal(cp.get("cj_bbox"), d["s0_cj_bbox"])
    np.testing.assert_array_equal(cp.get("ghost_map"), d["s0_ghost_map"])
    np.testing.assert_array_equal(cp.get("ghost_pbc"), d["s0_ghost_pbc"])
    ns = _nslots(d)
    np.testing.assert_array_equal(cp.get("slot_x")[:3 * ns], d["s0_slot_x"])
    np.testing.assert_array_equal(cp.get("slot_t")[:ns], d["s0_slot_t"])
    np.testing.assert_array_equal(cp.get("numneigh"), d["s0_numneigh"])
    np.testing.assert_array_equal(cp.get("numneigh_masked"), d["s0_numneigh_masked"])
    np.testing.assert_array_equal(cp.get("neighbors"), d["s0_neighbors"])
    np.testing.assert_array_equal(cp.get("imask"), d["s0_imask"].view(np.uint32))


@pytest.mark.parametrize("size", ["4", "4_half", "5x6x7"])
def test_m1n2_force_and_trajectory_bit_identical(size):
    d = load_golden("cp_m1n2_%s" % size)
    cp, half, args = _setup(d, "m1n2")
    cp.reneighbour(half)
    f, _, _ = cp.force(half)
    nci = meta(d, "s0_nci")
    np.testing.assert_array_equal(f[:3 * nci], d["s0_slot_f"])
    cp2, half, args = _setup(d, "m1n2")
    ks = cp2.run(half, args[4], args[5], dt=meta(d, "dt"), dtforce=meta(d, "dtforce"))
    np.testing.assert_array_equal(ks, d["traj_ksum"])
    np.testing.assert_array_equal(cp2.get("slot_x")[:3 * _nslots(d, "end")], d["end_slot_x"])


@pytest.mark.parametrize("variant,size", CASES)
def test_reneighbour_after_motion_matches_reference(variant, size):
    """The first rebuild after motion: from the atoms the reference handed to it (cluster order
    of the previous build, wrapped into the box), clusters, ghosts and lists must match the
    reference's rebuilt structures -- positions are off-lattice here, so no ties."""
    d = load_golden("cp_%s_%s" % (variant, size))
    args = [int(str(a)) for a in d["args"]]
    nx, ny, nz, half = args[:4]
    lat = meta(d, "lattice")
    M, N = SHAPES[variant]
    cp = ob.CP(M, N, (nx * lat, ny * lat, nz * lat), meta(d, "cutneigh"), d["pre1_x"],
               d["pre1_v"], d["pre1_type"])
    cp.reneighbour(half)
    assert cp.info()["nci"] == meta(d, "r1_nci")
    assert cp.info()["nghost"] == meta(d, "r1_nghost_clusters")
    np.testing.assert_array_equal(cp.get("ci_bbox"), d["r1_ci_bbox"])
    np.testing.assert_array_equal(cp.get("ghost_map"), d["r1_ghost_map"])
    np.testing.assert_array_equal(cp.get("slot_x")[:3 * _nslots(d, "r1")], d["r1_slot_x"])
    np.testing.assert_array_equal(cp.get("neighbors"), d["r1_neighbors"])
    np.testing.assert_array_equal(cp.get("numneigh_masked"), d["r1_numneigh_masked"])


def _csr_sets(numneigh, neighbors):
    off = np.concatenate([[0], np.cumsum(numneigh)])
    return [neighbors[off[c]:off[c + 1]] for c in range(numneigh.size)]


@pytest.mark.parametrize("variant", list(SHAPES))
def test_prune_matches_reference(variant):
    """pruneNeighborCPU (neighbor.c:1099-1264) on the first build's lists after the golden
    displacement of the local slots (ghosts refreshed by the reference): list contents, order,
    masked counts and masks equal the reference's.  Independently of the reference, every entry
    kept has an atom pair inside cutneigh and every entry dropped has none (brute force)."""
    d = load_golden("cp_%s_prune" % variant)
    cp, half, _ = _setup(d, variant)
    cp.reneighbour(half)
    np.testing.assert_array_equal(cp.get("neighbors"), d["s0_neighbors"])
    cp.set_slot_x(d["pr_slot_x"])
    removed = cp.prune()
    assert removed == d["s0_numneigh"].sum() - d["pr_numneigh"].sum() > 0
    np.testing.assert_array_equal(cp.get("numneigh"), d["pr_numneigh"])
    np.testing.assert_array_equal(cp.get("numneigh_masked"), d["pr_numneigh_masked"])
    np.testing.assert_array_equal(cp.get("neighbors"), d["pr_neighbors"])
    np.testing.assert_array_equal(cp.get("imask"), d["pr_imask"].view(np.uint32))
    M, N = SHAPES[variant]
    xs = d["pr_slot_x"].reshape(-1, 3)
    cut2 = meta(d, "cutneigh") ** 2
    before = _csr_sets(d["s0_numneigh"], d["s0_neighbors"])
    after = _csr_sets(d["pr_numneigh"], d["pr_neighbors"])
    for ci in range(len(before)):
        xi = xs[ci * M:(ci + 1) * M]
        kept = set(after[ci].tolist())
        assert kept <= set(before[ci].tolist())
        for cj in before[ci]:
            xj = xs[cj * N:(cj + 1) * N]
            r2 = ((xi[:, None, :] - xj[None, :, :]) ** 2).sum(-1).min()
            assert (r2 < cut2) == (cj in kept), (ci, cj, r2)


@pytest.mark.parametrize("variant", ["m4n4", "m4n8"])
@pytest.mark.parametrize("half", [0, 1])
def test_mxn_force_equals_verlet_force(variant, half):
    """On the same coordinates the cluster-pair force (any M x N, full or half list) and the
    Verlet-list force sum the same pairs within the cutoff: equal to FP64 rounding."""
    M, N = SHAPES[variant]
    d = load_golden("vl_lj_4")  # a moved state: end of a 40-step reference run
    nl = meta(d, "end_nlocal")
    x = d["end_x"][:3 * nl].copy()
    lat = meta(d, "lattice")
    prd = (4 * lat, 4 * lat, 4 * lat)
    ob.lib().orc_update_atoms_pbc(nl, x, *prd)
    v = np.zeros_like(x)
    t = np.zeros(nl, np.int32)
    cp = ob.CP(M, N, prd, 2.8, x, v, t)
    cp.reneighbour(half)
    f, e, w = cp.force(half)
    # map slot forces back to atoms: atom order inside clusters follows the cluster build
    slots = cp.get("slot_x").reshape(-1, 3)
    nci = cp.info()["nci"]
    nat = cp.get("ci_natoms")
    live = np.concatenate([np.arange(ci * M, ci * M + nat[ci]) for ci in range(nci)])
    xs, fs = slots[live], f.reshape(-1, 3)[live]
    # Verlet reference on the same atoms
    g = ob.vl_grid(4, 4, 4, lat, 2.8)
    xa = np.ascontiguousarray(xs.reshape(-1))
    bmap, pbc = ob.setup_pbc(nl, xa, prd, 2.8)
    xall = np.zeros(3 * (nl + bmap.size))
    xall[:3 * nl] = xa
    xall = ob.update_pbc(nl, xall, bmap, pbc, prd)
    tt = np.zeros(nl + bmap.size, np.int32)
    one = np.ones(16)
    nn, nb = ob.build_neighbor(g, nl, xall, tt, 4, one * 7.84, 0)
    fv, ev, wv = ob.force_lj(nl, xall, tt, 4, one, one, one * 6.25, nn, nb, 0)
    fv = fv[:3 * nl].reshape(-1, 3)
    assert np.linalg.norm(fs - fv) / np.linalg.norm(fv) < 1e-13
    assert abs(e - ev) <= 1e-12 * abs(ev)
    assert abs(w - wv) <= 1e-12 * abs(wv)
