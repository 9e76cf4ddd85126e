"""The drop-in structs (include/mdbench_vl.h, include/mdbench_cp.h) are layout-identical to
MD-Bench's own: a small C program prints sizeof/offsetof of every field, compiled once against
the reference's headers (with the -D set of its GCC build) and once against ours.  Runs only where
the reference sources are present (the build container); elsewhere it is skipped."""
import os
import subprocess
import tempfile

import pytest

from conftest import REPO

REF = os.environ.get("MDBENCH_REF", "/root/reference")
pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")),
                                reason="reference sources not present")

DEFS = ["-DATOM_POSITION_AOS", "-DNBLIST_AOS", "-DPRECISION=2", "-DLJ_COMB_RULE=1",
        "-DALIGNMENT=64", "-D_GNU_SOURCE"]

PARAM = ["force_field", "param_file", "input_file", "vtk_file", "xtc_file", "write_atom_file",
         "types_file", "epsilon", "sigma", "sigma6", "temp", "rho", "mass", "ntypes",
         "epsilon_per_type", "sigma_per_type", "ntimes", "nstat", "reneigh_every",
         "resort_every", "prune_every", "x_out_every", "v_out_every", "half_neigh", "dt",
         "dtforce", "skin", "cutforce", "cutneigh", "nx", "ny", "nz", "pbc_x", "pbc_y", "pbc_z",
         "lattice", "xlo", "xhi", "ylo", "yhi", "zlo", "zhi", "xprd", "yprd", "zprd",
         "proc_freq", "super_clustering", "eam_file", "balance", "method", "balance_every",
         "setup"]
VL = {
    "Parameter": PARAM,
    "Box": ["id", "xprd", "yprd", "zprd", "lo", "hi"],
    "Atom": ["Natoms", "Nlocal", "Nghost", "Nmax", "x", "y", "z", "vx", "vy", "vz", "fx", "fy",
             "fz", "border_map", "type", "ntypes", "epsilon", "sigma6", "cutforcesq",
             "cutneighsq", "sqrt_epsilon", "sigma3", "sqrt_epsilon_per_type", "sigma3_per_type",
             "d_atom", "d_atom.x", "d_atom.fx", "d_atom.border_map", "d_atom.sigma3", "mybox"],
    "Neighbor": ["every", "ncalls", "maxneighs", "half_neigh", "neighbors", "numneigh",
                 "d_neighbor", "d_neighbor.numneigh", "half_stencil", "Nshell",
                 "numNeighShell", "neighshell", "listshell"],
    "Stats": ["total_force_neighs", "total_force_iters", "atoms_within_cutoff",
              "atoms_outside_cutoff"],
}
CP = {
    "Parameter": PARAM,
    "Cluster": ["natoms", "bbminx", "bbmaxx", "bbminy", "bbmaxy", "bbminz", "bbmaxz"],
    "Atom": ["Natoms", "Nlocal", "Nghost", "Nmax", "Nclusters_local", "Nclusters_ghost",
             "Nclusters_max", "NmaxGhost", "ncj", "x", "y", "z", "vx", "vy", "vz",
             "border_map", "type", "ntypes", "epsilon", "sigma6", "cutforcesq", "cutneighsq",
             "sqrt_epsilon_per_type", "sigma3_per_type", "PBCx", "PBCy", "PBCz", "cl_x", "cl_v",
             "cl_f", "cl_t", "cl_sqrt_epsilon", "cl_sigma3", "iclusters", "jclusters",
             "siclusters", "cluster_bin", "dummy_cj", "exclusion_filter",
             "diagonal_4xn_j_minus_i", "diagonal_2xnn_j_minus_i", "masks_2xnn_hn",
             "masks_2xnn_fn", "masks_4xn_hn", "masks_4xn_fn", "mybox"],
    "Neighbor": ["every", "ncalls", "maxneighs", "numneigh", "numneigh_masked", "half_neigh",
                 "neighbors", "neighbors_imask", "Nshell", "numNeighShell", "neighshell",
                 "listshell"],
    "Stats": ["calculated_forces", "num_neighs", "force_iters", "atoms_within_cutoff",
              "atoms_outside_cutoff", "clusters_within_cutoff", "clusters_outside_cutoff"],
}


def _program(includes, structs):
    lines = ["#include <stdio.h>", "#include <stddef.h>"] + ["#include <%s>" % h for h in includes]
    lines.append("int main(void) {")
    for st, fields in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (st, st))
        for f in fields:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (st, f, st, f))
    lines.append("return 0; }")
    return "\n".join(lines)


def _layout(src, incs, extra):
    with tempfile.TemporaryDirectory() as wd:
        c = os.path.join(wd, "p.c")
        exe = os.path.join(wd, "p")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c11", "-w"] + DEFS + extra + ["-I" + i for i in incs] +
                       [c, "-o", exe], check=True)
        return subprocess.run([exe], capture_output=True, text=True, check=True).stdout


def test_verlet_list_structs_match_reference():
    ref = _layout(_program(["atom.h", "neighbor.h", "parameter.h", "stats.h"], VL),
                  [os.path.join(REF, "src", "verletlist"), os.path.join(REF, "src", "common")],
                  ["-DVECTOR_WIDTH=1"])
    ours = _layout(_program(["mdbench_vl.h"], VL), [os.path.join(REPO, "include")], [])
    assert ref == ours


def test_cluster_pair_structs_match_reference():
    ref = _layout(_program(["atom.h", "neighbor.h", "parameter.h", "stats.h"], CP),
                  [os.path.join(REF, "src", "clusterpair"), os.path.join(REF, "src", "common")],
                  ["-DVECTOR_WIDTH=2", "-DCLUSTER_PAIR", "-DUSE_REFERENCE_KERNEL",
                   "-DCLUSTERPAIR_KERNEL_AUTO"])
    ours = _layout(_program(["mdbench_cp.h"], CP), [os.path.join(REPO, "include")], [])
    assert ref == ours
