"""Write the golden fixtures of the Verlet-list path from the compiled reference.

Builds ``ref_dump_vl`` against ``oracle/_ref/libmdref_vl.a`` (MD-Bench's own Verlet-list and
common sources compiled by ``oracle/Makefile`` with strict IEEE flags), runs it on small seeded
cases and packs every blob it writes into ``tests/golden/<case>.npz``.  Run in a container where
``/root/reference`` exists:

    make -C oracle ref && python tests/golden/make_golden.py

The .npz files hold only data (inputs and the reference's outputs); the script that made them is
this file.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("MDBENCH_REF", "/root/reference")
REFLIB = os.path.join(REPO, "oracle", "_ref")
sys.path.insert(0, HERE)
from eam_tables import write_funcfl_synthetic  # noqa: E402

DEFS = ("-DATOM_POSITION_AOS -DNBLIST_AOS -DPOSITION_AOS3_SUP -DPRECISION=2 -DLJ_COMB_RULE=1 "
        "-DSUPERCLUSTER_INVERSE_THREAD_MAPPING -DALIGNMENT=64 -D_GNU_SOURCE -DNO_ZMM_INTRIN "
        "-DUSE_REFERENCE_KERNEL -DVECTOR_WIDTH=1").split()

DTYPES = {
    "x": np.float64, "v": np.float64, "f": np.float64, "spline": np.float64,
    "ksum": np.float64, "bbox": np.float64, "epsilon": np.float64, "sigma6": np.float64, "cutforcesq": np.float64,
    "cutneighsq": np.float64,
}


def _dtype(name):
    for suffix, dt in DTYPES.items():
        if name.endswith(suffix):
            return dt
    return np.int32


def build_dumper(workdir):
    exe = os.path.join(workdir, "ref_dump_vl")
    cmd = (["gcc", "-O2", "-std=c11", "-fopenmp", "-ffp-contract=off", "-fno-fast-math"] + DEFS +
           ["-I" + os.path.join(REF, "src", "verletlist"), "-I" + os.path.join(REF, "src", "common"),
            os.path.join(HERE, "ref_dump_vl.c"), os.path.join(REFLIB, "libmdref_vl.a"),
            "-lm", "-o", exe])
    subprocess.run(cmd, check=True)
    return exe


def run_case(exe, name, args, workdir):
    out = os.path.join(workdir, name)
    os.makedirs(out, exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS="1", OMP_SCHEDULE="static")
    subprocess.run([exe, out] + [str(a) for a in args], check=True, env=env,
                   stdout=subprocess.DEVNULL)
    data = {}
    with open(os.path.join(out, "meta.txt")) as fh:
        for line in fh:
            k, t, v = line.split()
            data["meta_" + k] = np.array(int(v) if t == "int" else float(v))
    for fn in sorted(os.listdir(out)):
        if fn.endswith(".bin"):
            key = fn[:-4]
            data[key] = np.fromfile(os.path.join(out, fn), dtype=_dtype(key))
    data["args"] = np.array([str(a) for a in args])
    dst = os.path.join(HERE, name + ".npz")
    np.savez_compressed(dst, **data)
    print("wrote", dst, sum(v.nbytes for v in data.values()), "bytes")


def main():
    with tempfile.TemporaryDirectory() as wd:
        exe = build_dumper(wd)
        types2 = os.path.join(wd, "types2.dat")
        with open(types2, "w") as fh:
            fh.write("# epsilon sigma\n1.0 1.0\n0.5 0.88\n")
        eam_syn = os.path.join(wd, "cu_synth.eam")
        write_funcfl_synthetic(eam_syn)
        # name, nx, ny, nz, half, nsteps, reneigh, kind, extra...
        run_case(exe, "vl_lj_4", [4, 4, 4, 0, 40, 20, "lj", "-"], wd)
        run_case(exe, "vl_lj_4_half", [4, 4, 4, 1, 40, 20, "lj", "-"], wd)
        run_case(exe, "vl_lj_5x6x7", [5, 6, 7, 0, 20, 10, "lj", "-"], wd)
        run_case(exe, "vl_lj_types", [4, 4, 4, 0, 20, 10, "lj", types2], wd)
        run_case(exe, "vl_eam_synth_4", [4, 4, 4, 0, 20, 10, "eam", eam_syn, 3.615, 1600.0], wd)
        run_case(exe, "vl_eam_cu_u3_4", [4, 4, 4, 0, 20, 10, "eam",
                                         os.path.join(REF, "data", "Cu_u3.eam"), 3.615, 600.0], wd)


if __name__ == "__main__" and os.environ.get("MDB_GOLDEN_VL", "1") == "1":
    main()


# ---------------------------------------------------------------- cluster pair
CP_VARIANTS = {
    # name: (archive, extra defines, march)
    "m1n2": ("libmdref_cp.a", "-DUSE_REFERENCE_KERNEL -DVECTOR_WIDTH=2", "-march=x86-64-v2"),
    "m4n4": ("libmdref_cp44.a", "-DVECTOR_WIDTH=4 -D__SIMD_KERNEL__ -D__ISA_AVX2__",
             "-march=x86-64-v3"),
    "m4n8": ("libmdref_cp48.a", "-DVECTOR_WIDTH=8 -D__SIMD_KERNEL__ -D__ISA_AVX512__",
             "-march=x86-64-v4"),
}
CP_BASE = ("-DATOM_POSITION_AOS -DNBLIST_AOS -DPOSITION_AOS3_SUP -DPRECISION=2 -DLJ_COMB_RULE=1 "
           "-DSUPERCLUSTER_INVERSE_THREAD_MAPPING -DALIGNMENT=64 -D_GNU_SOURCE -DNO_ZMM_INTRIN "
           "-DCLUSTER_PAIR -DCLUSTERPAIR_KERNEL_AUTO").split()


def build_cp_dumper(workdir, variant):
    arch, extra, march = CP_VARIANTS[variant]
    exe = os.path.join(workdir, "ref_dump_cp_" + variant)
    cmd = (["gcc", "-O2", "-std=c11", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", march]
           + CP_BASE + extra.split() +
           ["-I" + os.path.join(REF, "src", "clusterpair"), "-I" + os.path.join(REF, "src", "common"),
            os.path.join(HERE, "ref_dump_cp.c"), os.path.join(REFLIB, arch), "-lm", "-o", exe])
    subprocess.run(cmd, check=True)
    return exe


def main_cp():
    with tempfile.TemporaryDirectory() as wd:
        for variant in CP_VARIANTS:
            exe = build_cp_dumper(wd, variant)
            dump_force = 1 if variant == "m1n2" else 0
            for half in (0, 1):
                name = "cp_%s_4%s" % (variant, "_half" if half else "")
                run_case(exe, name, [4, 4, 4, half, 40, 20, dump_force], wd)
            run_case(exe, "cp_%s_5x6x7" % variant, [5, 6, 7, 0, 20, 10, dump_force], wd)
            # pruneNeighbor after a fixed displacement of the first build's slots
            run_case(exe, "cp_%s_prune" % variant, [4, 4, 4, 0, 0, 20, 0, 1], wd)


def main_cp_prune():
    with tempfile.TemporaryDirectory() as wd:
        for variant in CP_VARIANTS:
            exe = build_cp_dumper(wd, variant)
            run_case(exe, "cp_%s_prune" % variant, [4, 4, 4, 0, 0, 20, 0, 1], wd)


if __name__ == "__main__" and os.environ.get("MDB_GOLDEN_CP", "1") == "1":
    main_cp()
elif __name__ == "__main__" and os.environ.get("MDB_GOLDEN_CP_PRUNE", "0") == "1":
    main_cp_prune()
