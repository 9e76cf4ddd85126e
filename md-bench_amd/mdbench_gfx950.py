"""ctypes binding of libmdgfx950 (the C ABI in include/mdgfx950.h and include/mdbench_driver.h).

The library is built in-tree by ``make -C md-bench_amd``; importing this module never falls back
to anything else: a missing library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MDGFX950_LIB names another build of the same library (the sanitizer build of its host code)
LIB_PATH = os.environ.get("MDGFX950_LIB") or os.path.join(HERE, "lib", "libmdgfx950.so")

FF_LJ, FF_EAM = 0, 1

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Config(C.Structure):
    """mdb_config (include/mdbench_driver.h)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("force_field", C.c_int), ("half_neigh", C.c_int),
        ("ntimes", C.c_int), ("nstat", C.c_int), ("reneigh_every", C.c_int),
        ("rho", C.c_double), ("lattice_a", C.c_double), ("temp", C.c_double),
        ("dt", C.c_double), ("cutforce", C.c_double), ("skin", C.c_double),
        ("ntypes", C.c_int),
        ("epsilon_per_type", C.POINTER(C.c_double)), ("sigma_per_type", C.POINTER(C.c_double)),
        ("eam_file", C.c_char_p), ("strict", C.c_int),
        ("cluster_m", C.c_int), ("cluster_n", C.c_int), ("precision", C.c_int),
        ("compensated", C.c_int),
        ("pair_epsilon", C.POINTER(C.c_double)), ("pair_sigma", C.POINTER(C.c_double)),
        ("pair_cutoff", C.POINTER(C.c_double)), ("type1_fraction", C.c_double),
        ("type_seed", C.c_int), ("input_file", C.c_char_p),
        ("prune_every", C.c_int), ("resort_every", C.c_int), ("maxwell_seed", C.c_int),
    ]


class Box(C.Structure):
    _fields_ = [("id", C.c_int), ("xprd", C.c_double), ("yprd", C.c_double),
                ("zprd", C.c_double), ("lo", C.c_double * 3), ("hi", C.c_double * 3)]


_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)


class DeviceAtom(C.Structure):
    _fields_ = [(n, _PD) for n in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")] + \
               [("border_map", _PI), ("type", _PI)] + \
               [(n, _PD) for n in ("epsilon", "sigma6", "cutforcesq", "cutneighsq",
                                   "sqrt_epsilon", "sigma3")]


class Atom(C.Structure):
    """Atom of src/verletlist/atom.h:31-54 (include/mdbench_vl.h)."""
    _fields_ = [("Natoms", C.c_int), ("Nlocal", C.c_int), ("Nghost", C.c_int),
                ("Nmax", C.c_int)] + \
               [(n, _PD) for n in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")] + \
               [("border_map", _PI), ("type", _PI), ("ntypes", C.c_int)] + \
               [(n, _PD) for n in ("epsilon", "sigma6", "cutforcesq", "cutneighsq",
                                   "sqrt_epsilon", "sigma3", "sqrt_epsilon_per_type",
                                   "sigma3_per_type")] + \
               [("d_atom", DeviceAtom), ("mybox", Box)]


class Neighbor(C.Structure):
    """Neighbor of src/verletlist/neighbor.h:29-46 (include/mdbench_vl.h)."""
    _fields_ = [("every", C.c_int), ("ncalls", C.c_int), ("maxneighs", C.c_int),
                ("half_neigh", C.c_int), ("neighbors", _PI), ("numneigh", _PI),
                ("d_neighbors", _PI), ("d_numneigh", _PI), ("half_stencil", C.c_int),
                ("Nshell", C.c_int), ("numNeighShell", _PI), ("neighshell", _PI),
                ("listshell", _PI)]


class Parameter(C.Structure):
    """Parameter of src/common/parameter.h:76-119 (include/mdbench_vl.h)."""
    _fields_ = [("force_field", C.c_int)] + \
               [(n, C.c_char_p) for n in ("param_file", "input_file", "vtk_file", "xtc_file",
                                          "write_atom_file", "types_file")] + \
               [(n, C.c_double) for n in ("epsilon", "sigma", "sigma6", "temp", "rho", "mass")] + \
               [("ntypes", C.c_int), ("epsilon_per_type", _PD), ("sigma_per_type", _PD)] + \
               [(n, C.c_int) for n in ("ntimes", "nstat", "reneigh_every", "resort_every",
                                       "prune_every", "x_out_every", "v_out_every",
                                       "half_neigh")] + \
               [(n, C.c_double) for n in ("dt", "dtforce", "skin", "cutforce", "cutneigh")] + \
               [(n, C.c_int) for n in ("nx", "ny", "nz", "pbc_x", "pbc_y", "pbc_z")] + \
               [(n, C.c_double) for n in ("lattice", "xlo", "xhi", "ylo", "yhi", "zlo", "zhi",
                                          "xprd", "yprd", "zprd", "proc_freq")] + \
               [("super_clustering", C.c_int), ("eam_file", C.c_char_p)] + \
               [(n, C.c_int) for n in ("balance", "method", "balance_every", "setup")]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmdgfx950.so is not built (run `make -C md-bench_amd`): %s"
                           % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.mdbDefaultConfig.argtypes = [C.POINTER(Config)]
    L.mdbVLCreate.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.mdbVLRun.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.mdbVLDestroy.argtypes = [vp]
    for fn in ("mdbVLParam", "mdbVLAtom", "mdbVLNeighbor", "mdbVLStats"):
        getattr(L, fn).argtypes = [vp]
        getattr(L, fn).restype = vp
    L.mdbThermo.argtypes = [vp, C.c_double, _PD, _PD]
    L.mdbVLInitialState.argtypes = [vp, _dp, _dp, _ip]
    L.mdbCreateLattice.argtypes = [C.POINTER(Config), _dp, _dp, _ip, C.c_int]
    L.mdbMyrandom.argtypes = [C.POINTER(C.c_int)]
    L.mdbMyrandom.restype = C.c_double
    L.mdbWriteFuncflSynthetic.argtypes = [C.c_char_p]
    L.mdbStubRun.argtypes = [C.c_int] * 7 + [_dp]
    L.mdbWriteDmp.argtypes = [C.c_char_p, C.c_int, _dp, _dp, _ip, _dp, C.c_int]
    L.mdbReadDmp.argtypes = [C.c_char_p, C.c_int, _dp, _dp, _ip, _dp, C.POINTER(C.c_int)]
    L.mdgfx950Status.restype = C.c_int
    L.mdgfx950StatusMessage.restype = C.c_char_p
    L.mdgfx950SetStrict.argtypes = [C.c_int]
    L.mdgfx950EnableTiming.argtypes = [C.c_int]
    L.mdgfx950TimeOnly.argtypes = [C.c_char_p]
    L.mdgfx950TimeOnly.restype = C.c_int
    L.mdgfx950TimeEvery.argtypes = [C.c_int]
    L.mdgfx950TimeEvery.restype = None
    L.mdgfx950KernelTime.argtypes = [C.c_char_p, _PD, C.POINTER(C.c_longlong)]
    # drop-in entry points (opaque struct pointers from the driver)
    for fn in ("computeForceLJGFX950", "computeForceEamGFX950"):
        getattr(L, fn).argtypes = [vp, vp, vp, vp]
        getattr(L, fn).restype = C.c_double
    L.buildNeighborGFX950.argtypes = [vp, vp]
    L.binatomsGFX950.argtypes = [vp]
    L.sortAtomGFX950.argtypes = [vp]
    L.setupPbcGFX950.argtypes = [vp, vp]
    L.updatePbcGFX950.argtypes = [vp, vp, C.c_bool]
    L.updateAtomsPbcGFX950.argtypes = [vp, vp, C.c_bool]
    L.initialIntegrateGFX950.argtypes = [C.c_bool, vp, vp]
    L.finalIntegrateGFX950.argtypes = [C.c_bool, vp, vp]
    L.kineticSumGFX950.argtypes = [vp, vp]
    L.kineticSumGFX950.restype = C.c_double
    L.requestEnergyVirialGFX950.argtypes = [C.c_int]
    L.requestEamAtomEnergyGFX950.argtypes = [C.c_int]
    L.downloadEamAtomEnergyGFX950.argtypes = [vp, vp]
    L.downloadEamAtomEnergyGFX950.restype = C.c_int
    L.energyVirialGFX950.argtypes = [_PD, _PD]
    L.copyDataToGFX950Device.argtypes = [vp, vp]
    L.copyDataFromGFX950Device.argtypes = [vp, vp]
    L.downloadVLStateGFX950.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_longlong]
    L.downloadBinsGFX950.argtypes = [vp, C.c_int, vp, C.c_int]
    # cluster-pair scheme
    L.mdbCPCreate.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.mdbCPRun.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.mdbCPSetAtoms.argtypes = [vp, _dp, _dp, _ip, C.c_int]
    L.mdbCPDestroy.argtypes = [vp]
    for fn in ("mdbCPParam", "mdbCPAtom", "mdbCPNeighbor", "mdbCPStats"):
        getattr(L, fn).argtypes = [vp]
        getattr(L, fn).restype = vp
    L.mdbCPThermo.argtypes = [vp, C.c_double, _PD, _PD]
    L.computeForceLJClusterGFX950.argtypes = [vp, vp, vp, vp]
    L.computeForceLJClusterGFX950.restype = C.c_double
    L.kineticSumClusterGFX950.argtypes = [vp, vp]
    L.kineticSumClusterGFX950.restype = C.c_double
    L.kineticSumClusterNoAllocGFX950.argtypes = [vp, vp, vp]
    L.kineticSumClusterNoAllocGFX950.restype = C.c_int
    L.requestEnergyVirialClusterGFX950.argtypes = [C.c_int]
    L.energyVirialClusterGFX950.argtypes = [_PD, _PD]
    L.downloadClusterArrayGFX950.argtypes = [C.c_char_p, vp, C.c_longlong]
    L.downloadClusterArrayGFX950.restype = C.c_longlong
    L.uploadClusterArrayGFX950.argtypes = [C.c_char_p, vp, C.c_longlong]
    L.uploadClusterArrayGFX950.restype = C.c_longlong
    L.pruneNeighborClusterGFX950.argtypes = [vp, vp, vp]
    L.pruneNeighborClusterGFX950.restype = None
    for fn in ("initialIntegrateClusterGFX950", "finalIntegrateClusterGFX950"):
        getattr(L, fn).argtypes = [vp, vp]
        getattr(L, fn).restype = None
    L.updatePbcClusterGFX950.argtypes = [vp, vp, C.c_bool]
    L.updatePbcClusterGFX950.restype = None
    L.computeForceFinalClusterGFX950.argtypes = [vp, vp, vp, vp]
    L.computeForceFinalClusterGFX950.restype = C.c_double
    L.computeForceFinalInitialClusterGFX950.argtypes = [vp, vp, vp, vp]
    L.computeForceFinalInitialClusterGFX950.restype = C.c_double
    L.computeForceFinalInitialGFX950.argtypes = [vp, vp, vp, vp]
    L.computeForceFinalInitialGFX950.restype = C.c_double
    L.clusterInfoGFX950.argtypes = [_ip]
    L.setClusterPrecisionGFX950.argtypes = [C.c_int, C.c_int]
    _lib = L
    return L


def status():
    L = lib()
    return L.mdgfx950Status(), (L.mdgfx950StatusMessage() or b"").decode()


def default_config(**kw):
    cfg = Config()
    lib().mdbDefaultConfig(C.byref(cfg))
    keep = []
    for k, v in kw.items():
        if k in ("epsilon_per_type", "sigma_per_type", "pair_epsilon", "pair_sigma",
                 "pair_cutoff") and v is not None:
            arr = (C.c_double * len(v))(*v)
            keep.append(arr)
            setattr(cfg, k, C.cast(arr, C.POINTER(C.c_double)))
        elif k in ("eam_file", "input_file"):
            setattr(cfg, k, v.encode() if v else None)
        else:
            setattr(cfg, k, v)
    cfg._keep = keep
    return cfg


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class VLSim:
    """A Verlet-list simulation living on the GPU (driver API of include/mdbench_driver.h)."""

    def __init__(self, **kw):
        self.cfg = default_config(**kw)
        h = C.c_void_p()
        rc = lib().mdbVLCreate(C.byref(self.cfg), C.byref(h))
        if rc:
            raise RuntimeError("mdbVLCreate failed (%d): %s" % (rc, status()[1]))
        self.h = h
        L = lib()
        self.param_p = L.mdbVLParam(h)
        self.atom_p = L.mdbVLAtom(h)
        self.neighbor_p = L.mdbVLNeighbor(h)
        self.stats_p = L.mdbVLStats(h)

    # struct views
    @property
    def param(self):
        return Parameter.from_address(self.param_p)

    @property
    def atom(self):
        return Atom.from_address(self.atom_p)

    @property
    def neighbor(self):
        return Neighbor.from_address(self.neighbor_p)

    def close(self):
        if getattr(self, "h", None):
            lib().mdbVLDestroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, nsteps, trace=False):
        ks = np.zeros(nsteps) if trace else None
        rc = lib().mdbVLRun(self.h, nsteps, ks.ctypes.data_as(_PD) if trace else None)
        if rc:
            raise RuntimeError("mdbVLRun failed (%d): %s" % (rc, status()[1]))
        return ks

    def thermo(self, ksum):
        t, p = C.c_double(), C.c_double()
        lib().mdbThermo(self.h, ksum, C.byref(t), C.byref(p))
        return t.value, p.value

    def kinetic_sum(self):
        return lib().kineticSumGFX950(self.param_p, self.atom_p)

    def initial_state(self):
        n = self.atom.Nlocal
        x, v, t = np.zeros(3 * n), np.zeros(3 * n), np.zeros(n, np.int32)
        lib().mdbVLInitialState(self.h, x, v, t)
        return x, v, t

    def types(self):
        """The local atoms' types in the device's current order (copyDataFromGFX950Device)."""
        L = lib()
        L.copyDataFromGFX950Device(self.param_p, self.atom_p)
        n = self.atom.Nlocal
        return np.ctypeslib.as_array(self.atom.type, (n,)).copy()

    def state(self, lists=True):
        """Download x (local+ghost), f, v, ghost map and the neighbour list (CSR)."""
        a = self.atom
        nl, ng = a.Nlocal, a.Nghost
        x = np.zeros(3 * (nl + ng))
        f = np.zeros(3 * nl)
        v = np.zeros(3 * nl)
        bmap = np.zeros(max(ng, 1), np.int32)
        pbc = np.zeros(3 * max(ng, 1), np.int32)
        out = dict(x=x, f=f, v=v, border_map=bmap[:ng], pbc=pbc[:3 * ng], nlocal=nl, nghost=ng)
        if lists:
            cap = nl * self.neighbor.maxneighs
            numneigh = np.zeros(nl, np.int32)
            csr = np.zeros(max(cap, 1), np.int32)
            rc = lib().downloadVLStateGFX950(self.atom_p, self.neighbor_p, _ptr(x), _ptr(f),
                                             _ptr(v), _ptr(bmap), _ptr(pbc), _ptr(numneigh),
                                             _ptr(csr), cap)
            assert rc >= 0, (rc, status())
            out.update(numneigh=numneigh, neighbors=csr[:rc].copy())
        else:
            rc = lib().downloadVLStateGFX950(self.atom_p, self.neighbor_p, _ptr(x), _ptr(f),
                                             _ptr(v), _ptr(bmap), _ptr(pbc), None, None, 0)
            assert rc >= 0, (rc, status())
        return out

    def bins(self, mbins_cap=1 << 22):
        nall = self.atom.Nlocal + self.atom.Nghost
        start = np.zeros(mbins_cap, np.int32)
        atoms = np.zeros(nall + 1, np.int32)
        mb = lib().downloadBinsGFX950(_ptr(start), mbins_cap, _ptr(atoms), nall + 1)
        assert mb > 0, (mb, status())
        return start[:mb + 1].copy(), atoms[:nall].copy()

    # drop-in entry points on this simulation's structs
    def compute_force(self):
        L = lib()
        fn = L.computeForceEamGFX950 if self.param.force_field == FF_EAM else \
            L.computeForceLJGFX950
        return fn(self.param_p, self.atom_p, self.neighbor_p, self.stats_p)

    def energy_virial(self):
        e, w = C.c_double(), C.c_double()
        lib().energyVirialGFX950(C.byref(e), C.byref(w))
        return e.value, w.value

    def eam_atom_energy(self):
        """Forces and per-atom energies F(rho_i) + 1/2 sum phi of the current positions."""
        L = lib()
        L.requestEamAtomEnergyGFX950(1)
        try:
            self.compute_force()
            out = np.zeros(self.atom.Nlocal)
            rc = L.downloadEamAtomEnergyGFX950(self.atom_p, _ptr(out))
            assert rc == 0, status()
        finally:
            L.requestEamAtomEnergyGFX950(0)
        return out


class CPSim:
    """A cluster-pair simulation living on the GPU (include/mdbench_driver_cp.h)."""

    INFO = ["nci", "ncj", "nghost", "dummy_cj", "nghost_atoms", "mbins", "nstencil",
            "maxneighs", "M", "N", "natoms"]
    F64 = ("slot_x", "slot_v", "slot_f", "ci_bbox", "cj_bbox", "atoms_x", "atoms_v")

    def __init__(self, **kw):
        self.cfg = default_config(**kw)
        h = C.c_void_p()
        rc = lib().mdbCPCreate(C.byref(self.cfg), C.byref(h))
        if rc:
            raise RuntimeError("mdbCPCreate failed (%d): %s" % (rc, status()[1]))
        self.h = h
        L = lib()
        self.param_p = L.mdbCPParam(h)
        self.atom_p = L.mdbCPAtom(h)
        self.neighbor_p = L.mdbCPNeighbor(h)
        self.stats_p = L.mdbCPStats(h)

    def close(self):
        if getattr(self, "h", None):
            lib().mdbCPDestroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, nsteps, trace=False):
        ks = np.zeros(nsteps) if trace else None
        rc = lib().mdbCPRun(self.h, nsteps, ks.ctypes.data_as(_PD) if trace else None)
        if rc:
            raise RuntimeError("mdbCPRun failed (%d): %s" % (rc, status()[1]))
        return ks

    def set_atoms(self, x, v, t):
        rc = lib().mdbCPSetAtoms(self.h, np.ascontiguousarray(x, np.float64),
                                 np.ascontiguousarray(v, np.float64),
                                 np.ascontiguousarray(t, np.int32), t.size)
        if rc:
            raise RuntimeError("mdbCPSetAtoms failed (%d): %s" % (rc, status()[1]))

    def info(self):
        arr = np.zeros(11, np.int32)
        lib().clusterInfoGFX950(arr)
        return dict(zip(self.INFO, arr.tolist()))

    def get(self, name):
        dt = np.float64 if name in self.F64 else np.int32
        cap = 1 << 25
        buf = np.zeros(cap, dt)
        n = lib().downloadClusterArrayGFX950(name.encode(), buf.ctypes.data_as(C.c_void_p), cap)
        if n < 0:
            raise RuntimeError("download %s failed (%d): %s" % (name, n, status()[1]))
        return buf[:n].copy()

    def set(self, name, arr):
        """Overwrites a device array from host (slot_x, slot_v)."""
        a = np.ascontiguousarray(arr, np.float64)
        n = lib().uploadClusterArrayGFX950(name.encode(), a.ctypes.data_as(C.c_void_p), a.size)
        if n != a.size:
            raise RuntimeError("upload %s failed (%d): %s" % (name, n, status()[1]))

    def prune(self):
        """pruneNeighbor on the current lists and slot positions."""
        lib().pruneNeighborClusterGFX950(self.param_p, self.atom_p, self.neighbor_p)
        st = status()
        if st[0]:
            raise RuntimeError("prune failed: %s" % st[1])

    def compute_force(self):
        return lib().computeForceLJClusterGFX950(self.param_p, self.atom_p, self.neighbor_p,
                                                 self.stats_p)

    def kinetic_sum(self):
        return lib().kineticSumClusterGFX950(self.param_p, self.atom_p)

    def thermo(self, ksum):
        t, p = C.c_double(), C.c_double()
        lib().mdbCPThermo(self.h, ksum, C.byref(t), C.byref(p))
        return t.value, p.value

    def energy_virial(self):
        lib().requestEnergyVirialClusterGFX950(1)
        self.compute_force()
        e, w = C.c_double(), C.c_double()
        lib().energyVirialClusterGFX950(C.byref(e), C.byref(w))
        lib().requestEnergyVirialClusterGFX950(0)
        return e.value, w.value


def create_lattice(**kw):
    cfg = default_config(**kw)
    n = 4 * cfg.nx * cfg.ny * cfg.nz
    x, v, t = np.zeros(3 * n), np.zeros(3 * n), np.zeros(n, np.int32)
    rc = lib().mdbCreateLattice(C.byref(cfg), x, v, t, n)
    assert rc >= 0, rc
    return x[:3 * rc], v[:3 * rc], t[:rc]


STUB_PATTERNS = {"seq": 0, "fix": 1, "rand": 2}


def stub_run(pattern, natoms, nneighs=76, nreps=1, ntimes=200, half=0, strict=0):
    """MD-Bench's synthetic-pattern force benchmark on the device (include/mdbench_driver.h)."""
    out = np.zeros(5)
    rc = lib().mdbStubRun(STUB_PATTERNS[pattern], natoms, nneighs, nreps, ntimes, half, strict,
                          out)
    if rc:
        raise RuntimeError("mdbStubRun failed (%d): %s" % (rc, status()[1]))
    return dict(seconds=out[0], atom_updates_per_s=out[1], ms_per_call=out[2], f0=out[3],
                flast=out[4])


def write_dmp(path, x, v, types, box, timestep=0):
    """Atoms in MD-Bench's dump format (box = xlo xhi ylo yhi zlo zhi)."""
    x = np.ascontiguousarray(x, np.float64)
    rc = lib().mdbWriteDmp(path.encode(), x.size // 3, x, np.ascontiguousarray(v, np.float64),
                           np.ascontiguousarray(types, np.int32),
                           np.ascontiguousarray(box, np.float64), timestep)
    assert rc == 0, rc


def read_dmp(path, cap):
    x, v, t = np.zeros(3 * cap), np.zeros(3 * cap), np.zeros(cap, np.int32)
    box = np.zeros(6)
    nt = C.c_int()
    n = lib().mdbReadDmp(path.encode(), cap, x, v, t, box, C.byref(nt))
    assert n > 0, n
    return x[:3 * n], v[:3 * n], t[:n], box, nt.value


def gpu_available():
    try:
        import torch  # noqa: F401  (device count only; no HIP context is created here)
        return torch.cuda.device_count() > 0
    except Exception:
        return False
