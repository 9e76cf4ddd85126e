"""ctypes binding of the CPU oracle (oracle/liboracle_md.so) -- test infrastructure only."""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# ORACLE_MD_LIB names another build of the oracle (its sanitizer build)
LIB_PATH = os.environ.get("ORACLE_MD_LIB") or os.path.join(REPO, "oracle", "liboracle_md.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class VLGrid(C.Structure):
    _fields_ = [
        ("xprd", C.c_double), ("yprd", C.c_double), ("zprd", C.c_double),
        ("cutneigh", C.c_double), ("cutneighsq", C.c_double),
        ("nbinx", C.c_int), ("nbiny", C.c_int), ("nbinz", C.c_int),
        ("binsizex", C.c_double), ("binsizey", C.c_double), ("binsizez", C.c_double),
        ("bininvx", C.c_double), ("bininvy", C.c_double), ("bininvz", C.c_double),
        ("pad_x", C.c_int), ("pad_y", C.c_int), ("pad_z", C.c_int),
        ("mbinx", C.c_int), ("mbiny", C.c_int), ("mbinz", C.c_int), ("mbins", C.c_int),
        ("nstencil", C.c_int), ("stencil", C.c_int * 1024),
    ]


class Funcfl(C.Structure):
    _fields_ = [("nrho", C.c_int), ("nr", C.c_int), ("drho", C.c_double), ("dr", C.c_double),
                ("cut", C.c_double), ("mass", C.c_double), ("frho", C.POINTER(C.c_double)),
                ("rhor", C.POINTER(C.c_double)), ("zr", C.POINTER(C.c_double))]


class Eam(C.Structure):
    _fields_ = [("nrho", C.c_int), ("nr", C.c_int), ("nrho_tot", C.c_int), ("nr_tot", C.c_int),
                ("dr", C.c_double), ("rdr", C.c_double), ("drho", C.c_double),
                ("rdrho", C.c_double), ("frho_spline", C.POINTER(C.c_double)),
                ("rhor_spline", C.POINTER(C.c_double)), ("z2r_spline", C.POINTER(C.c_double))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        _lib = load(LIB_PATH)
    return _lib


def load(path):
    """Loads an oracle build (the default one, or a copy built elsewhere) with its signatures."""
    L = C.CDLL(path)
    L.orc_vl_setup_grid.argtypes = [C.POINTER(VLGrid), C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_double]
    L.orc_vl_coord2bin.argtypes = [C.POINTER(VLGrid), C.c_double, C.c_double, C.c_double]
    L.orc_vl_binatoms.argtypes = [C.POINTER(VLGrid), C.c_int, _dp, _ip, _ip]
    L.orc_vl_build_neighbor.argtypes = [C.POINTER(VLGrid), C.c_int, C.c_int, _dp, _ip, C.c_int,
                                        _dp, C.c_int, _ip, _ip, C.c_longlong]
    L.orc_vl_build_neighbor.restype = C.c_longlong
    L.orc_vl_setup_pbc.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double,
                                   C.c_double, _ip, _ip, C.c_int]
    L.orc_vl_update_pbc.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, C.c_double, C.c_double,
                                    C.c_double]
    L.orc_update_atoms_pbc.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double]
    L.orc_vl_force_lj.argtypes = [C.c_int, C.c_int, _dp, _ip, C.c_int, _dp, _dp, _dp, _ip, _ip,
                                  C.c_int, _dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_eam_read_funcfl.argtypes = [C.c_char_p, C.POINTER(Funcfl)]
    L.orc_eam_setup.argtypes = [C.POINTER(Funcfl), C.POINTER(Eam)]
    L.orc_eam_free.argtypes = [C.POINTER(Funcfl), C.POINTER(Eam)]
    L.orc_vl_force_eam.argtypes = [C.POINTER(Eam), C.c_int, C.c_int, _dp, _ip, C.c_double, _ip,
                                   _ip, _dp, _dp, C.POINTER(C.c_double)]
    L.orc_vl_force_eam_atom.argtypes = [C.POINTER(Eam), C.c_int, C.c_int, _dp, _ip, C.c_double,
                                        _ip, _ip, _dp, _dp, _dp]
    L.orc_initial_integrate.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double]
    L.orc_final_integrate.argtypes = [C.c_int, _dp, _dp, C.c_double]
    L.orc_kinetic_sum.argtypes = [C.c_int, _dp, C.c_double]
    L.orc_kinetic_sum.restype = C.c_double
    L.orc_vl_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp, _ip,
                             C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int,
                             C.POINTER(Eam), C.c_double, C.c_double, C.c_int, C.c_int,
                             C.c_double, _dp, _dp]
    L.orc_vl_run_box.argtypes = [_dp, C.c_int, _dp, _dp, _ip, C.c_int, _dp, _dp, _dp, _dp,
                                 C.c_double, C.c_int, C.c_int, C.POINTER(Eam), C.c_double,
                                 C.c_double, C.c_int, C.c_int, C.c_double, _dp, _dp]
    L.orc_vl_sort_atoms.argtypes = [C.POINTER(VLGrid), C.c_int, _dp, _dp, _dp, _ip]
    L.orc_vl_run_resort.argtypes = L.orc_vl_run.argtypes + [C.c_int, _ip]
    return L


def vl_grid(nx, ny, nz, lattice, cutneigh):
    g = VLGrid()
    rc = lib().orc_vl_setup_grid(C.byref(g), nx, ny, nz, lattice, cutneigh)
    assert rc == 0
    return g


def binatoms(g, x):
    nall = x.size // 3
    start = np.zeros(g.mbins + 1, np.int32)
    atoms = np.zeros(nall + 1, np.int32)
    lib().orc_vl_binatoms(C.byref(g), nall, np.ascontiguousarray(x, np.float64), start, atoms)
    return start, atoms[:nall]


def build_neighbor(g, nlocal, x, type_, ntypes, cutneighsq, half, cap_per_atom=400):
    nall = x.size // 3
    numneigh = np.zeros(nlocal, np.int32)
    cap = nlocal * cap_per_atom
    nb = np.zeros(cap, np.int32)
    tot = lib().orc_vl_build_neighbor(C.byref(g), nlocal, nall, np.ascontiguousarray(x),
                                      np.ascontiguousarray(type_, np.int32), ntypes,
                                      np.ascontiguousarray(cutneighsq, np.float64), int(half),
                                      numneigh, nb, cap)
    assert tot >= 0
    return numneigh, nb[:tot].copy()


def setup_pbc(nlocal, x, prd, cutneigh, cap=None):
    cap = cap or 27 * nlocal + 1024
    bmap = np.zeros(cap, np.int32)
    pbc = np.zeros(3 * cap, np.int32)
    ng = lib().orc_vl_setup_pbc(nlocal, np.ascontiguousarray(x, np.float64), prd[0], prd[1],
                                prd[2], cutneigh, bmap, pbc, cap)
    assert ng >= 0
    return bmap[:ng].copy(), pbc[:3 * ng].copy()


def update_pbc(nlocal, x, bmap, pbc, prd):
    xx = np.ascontiguousarray(x, np.float64).copy()
    lib().orc_vl_update_pbc(nlocal, bmap.size, xx, bmap, pbc, prd[0], prd[1], prd[2])
    return xx


def force_lj(nlocal, x, type_, ntypes, eps, sig6, cutf, numneigh, nb, half, L=None):
    """Oracle LJ forces, pair energy and virial (L: an oracle build other than the default)."""
    nall = x.size // 3
    f = np.zeros(3 * nall, np.float64)
    e = C.c_double()
    w = C.c_double()
    (L or lib()).orc_vl_force_lj(nlocal, nall, np.ascontiguousarray(x, np.float64),
                          np.ascontiguousarray(type_, np.int32), ntypes,
                          np.ascontiguousarray(eps, np.float64),
                          np.ascontiguousarray(sig6, np.float64),
                          np.ascontiguousarray(cutf, np.float64), numneigh, nb, int(half), f,
                          C.byref(e), C.byref(w))
    return f, e.value, w.value


class EamTables:
    def __init__(self, path):
        self.file = Funcfl()
        self.eam = Eam()
        rc = lib().orc_eam_read_funcfl(path.encode(), C.byref(self.file))
        assert rc == 0, rc
        lib().orc_eam_setup(C.byref(self.file), C.byref(self.eam))

    def spline(self, which):
        e = self.eam
        n = e.nrho_tot if which == "frho" else e.nr_tot
        ptr = getattr(e, which + "_spline")
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def __del__(self):
        try:
            lib().orc_eam_free(C.byref(self.file), C.byref(self.eam))
        except Exception:
            pass


def force_eam(tab, nlocal, x, bmap, cutforcesq, numneigh, nb):
    nall = x.size // 3
    fp = np.zeros(nall, np.float64)
    f = np.zeros(3 * nlocal, np.float64)
    e = C.c_double()
    lib().orc_vl_force_eam(C.byref(tab.eam), nlocal, nall - nlocal,
                           np.ascontiguousarray(x, np.float64), bmap, cutforcesq, numneigh, nb,
                           fp, f, C.byref(e))
    return f, fp, e.value


def force_eam_atom(tab, nlocal, x, bmap, cutforcesq, numneigh, nb):
    """EAM forces and per-atom energies F(rho_i) + 1/2 sum_j phi(r_ij)."""
    nall = x.size // 3
    fp = np.zeros(nall, np.float64)
    f = np.zeros(3 * nlocal, np.float64)
    ea = np.zeros(nlocal, np.float64)
    lib().orc_vl_force_eam_atom(C.byref(tab.eam), nlocal, nall - nlocal,
                                np.ascontiguousarray(x, np.float64), bmap, cutforcesq, numneigh,
                                nb, fp, f, ea)
    return f, ea


def kinetic_sum(v, mass):
    v = np.ascontiguousarray(v, np.float64)
    return lib().orc_kinetic_sum(v.size // 3, v, mass)


def vl_run(nx, ny, nz, lattice, x, v, type_, ntypes, eps, sig6, cutf, cutn, cutneigh, half,
           ntimes, reneigh, dt, dtforce, mass, eam=None):
    nlocal = x.size // 3
    xx = np.ascontiguousarray(x, np.float64).copy()
    vv = np.ascontiguousarray(v, np.float64).copy()
    f = np.zeros(3 * nlocal)
    ks = np.zeros(ntimes + 1)
    rc = lib().orc_vl_run(nx, ny, nz, lattice, nlocal, xx, vv,
                          np.ascontiguousarray(type_, np.int32), ntypes,
                          np.ascontiguousarray(eps, np.float64),
                          np.ascontiguousarray(sig6, np.float64),
                          np.ascontiguousarray(cutf, np.float64),
                          np.ascontiguousarray(cutn, np.float64), cutneigh, int(half),
                          1 if eam is not None else 0,
                          C.byref(eam.eam) if eam is not None else None,
                          dt, dtforce, ntimes, reneigh, mass, f, ks)
    assert rc == 0, rc
    return xx, vv, f, ks


def vl_sort_atoms(g, x, v, f, type_):
    """sortAtom on the local atoms: the four arrays come back in bin order."""
    x, v, f = (np.ascontiguousarray(a, np.float64).copy() for a in (x, v, f))
    t = np.ascontiguousarray(type_, np.int32).copy()
    lib().orc_vl_sort_atoms(C.byref(g), x.size // 3, x, v, f, t)
    return x, v, f, t


def vl_run_resort(nx, ny, nz, lattice, x, v, type_, ntypes, eps, sig6, cutf, cutn, cutneigh,
                  half, ntimes, reneigh, dt, dtforce, mass, resort, eam=None):
    """vl_run with SORT_ATOMS every `resort` steps; also returns the final types."""
    nlocal = x.size // 3
    xx = np.ascontiguousarray(x, np.float64).copy()
    vv = np.ascontiguousarray(v, np.float64).copy()
    f = np.zeros(3 * nlocal)
    ks = np.zeros(ntimes + 1)
    tt = np.zeros(nlocal, np.int32)
    rc = lib().orc_vl_run_resort(nx, ny, nz, lattice, nlocal, xx, vv,
                                 np.ascontiguousarray(type_, np.int32), ntypes,
                                 np.ascontiguousarray(eps, np.float64),
                                 np.ascontiguousarray(sig6, np.float64),
                                 np.ascontiguousarray(cutf, np.float64),
                                 np.ascontiguousarray(cutn, np.float64), cutneigh, int(half),
                                 1 if eam is not None else 0,
                                 C.byref(eam.eam) if eam is not None else None,
                                 dt, dtforce, ntimes, reneigh, mass, f, ks, resort, tt)
    assert rc == 0, rc
    return xx, vv, f, ks, tt


def vl_run_box(box, x, v, type_, ntypes, eps, sig6, cutf, cutn, cutneigh, half, ntimes,
               reneigh, dt, dtforce, mass):
    """The Verlet-list run of atoms read from a file (bins sized from the box)."""
    nlocal = x.size // 3
    xx = np.ascontiguousarray(x, np.float64).copy()
    vv = np.ascontiguousarray(v, np.float64).copy()
    f = np.zeros(3 * nlocal)
    ks = np.zeros(ntimes + 1)
    rc = lib().orc_vl_run_box(np.ascontiguousarray(box, np.float64), nlocal, xx, vv,
                              np.ascontiguousarray(type_, np.int32), ntypes,
                              np.ascontiguousarray(eps, np.float64),
                              np.ascontiguousarray(sig6, np.float64),
                              np.ascontiguousarray(cutf, np.float64),
                              np.ascontiguousarray(cutn, np.float64), cutneigh, int(half), 0, None,
                              dt, dtforce, ntimes, reneigh, mass, f, ks)
    assert rc == 0, rc
    return xx, vv, f, ks


# ---------------------------------------------------------------- cluster pair
CP_RULES = {(1, 2): (2, 0), (4, 4): (4, 1), (4, 8): (8, 1)}  # (imask rule, fma in list test)


class CP:
    """Cluster-pair oracle context (oracle/oracle_cp.c)."""

    INFO = ["M", "N", "nci", "ncj", "nghost", "dummy_cj", "mbins", "nstencil", "nbinx", "nbiny",
            "mbinx", "mbiny", "mbinxlo", "mbinylo", "natoms"]

    def __init__(self, M, N, prd, cutneigh, x, v, t, fma_dist=None):
        L = lib()
        if not getattr(L, "_cp_bound", False):
            L.orc_cp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                        _ip]
            L.orc_cp_create.restype = C.c_void_p
            for fn in ("orc_cp_free", "orc_cp_build_clusters", "orc_cp_define_jclusters",
                       "orc_cp_setup_pbc", "orc_cp_bin_jclusters", "orc_cp_update_atoms"):
                getattr(L, fn).argtypes = [C.c_void_p]
            L.orc_cp_update_pbc.argtypes = [C.c_void_p, C.c_int]
            L.orc_cp_build_neighbor.argtypes = [C.c_void_p, C.c_int]
            L.orc_cp_build_neighbor.restype = C.c_longlong
            L.orc_cp_reneighbour.argtypes = [C.c_void_p, C.c_int]
            L.orc_cp_force_lj.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                          C.c_double, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]
            L.orc_cp_initial_integrate.argtypes = [C.c_void_p, C.c_double, C.c_double]
            L.orc_cp_final_integrate.argtypes = [C.c_void_p, C.c_double]
            L.orc_cp_kinetic_sum.argtypes = [C.c_void_p, C.c_double]
            L.orc_cp_kinetic_sum.restype = C.c_double
            L.orc_cp_run.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, _dp]
            L.orc_cp_run_prune.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                           C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                           C.c_int, C.c_double, _dp]
            L.orc_cp_prune.argtypes = [C.c_void_p]
            L.orc_cp_prune.restype = C.c_longlong
            L.orc_cp_info.argtypes = [C.c_void_p, _ip, C.POINTER(C.c_longlong)]
            L.orc_cp_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_longlong]
            L.orc_cp_get.restype = C.c_longlong
            L.orc_cp_set_slot_x.argtypes = [C.c_void_p, _dp, C.c_int]
            L._cp_bound = True
        rule, fm = CP_RULES[(M, N)]
        if fma_dist is not None:
            fm = fma_dist
        self.L = L
        self.h = L.orc_cp_create(M, N, rule, fm, prd[0], prd[1], prd[2], cutneigh,
                                 x.size // 3, np.ascontiguousarray(x, np.float64),
                                 np.ascontiguousarray(v, np.float64),
                                 np.ascontiguousarray(t, np.int32))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_cp_free(self.h)
            self.h = None

    def info(self):
        arr = np.zeros(15, np.int32)
        tot = C.c_longlong()
        self.L.orc_cp_info(self.h, arr, C.byref(tot))
        d = dict(zip(self.INFO, arr.tolist()))
        d["nb_total"] = tot.value
        return d

    def get(self, name):
        dt = np.float64 if name in ("slot_x", "slot_v", "slot_f", "ci_bbox", "cj_bbox",
                                    "atoms_x", "atoms_v") else np.int32
        if name == "imask":
            dt = np.uint32
        n = self.L.orc_cp_get(self.h, name.encode(), None, 0)
        n = -n if n < 0 else n
        # ask for the size: call with cap 0 returns -2 when non-empty
        cap = 1 << 26
        buf = np.zeros(cap if dt != np.float64 else cap // 2, dt)
        k = self.L.orc_cp_get(self.h, name.encode(), buf.ctypes.data_as(C.c_void_p), buf.size)
        assert k >= 0, (name, k)
        return buf[:k].copy()

    def reneighbour(self, half):
        rc = self.L.orc_cp_reneighbour(self.h, int(half))
        assert rc == 0, rc

    def force(self, half, eps=1.0, sig6=1.0, cutf=6.25):
        e, w = C.c_double(), C.c_double()
        self.L.orc_cp_force_lj(self.h, int(half), eps, sig6, cutf, C.byref(e), C.byref(w))
        return self.get("slot_f"), e.value, w.value

    def set_slot_x(self, x):
        x = np.ascontiguousarray(x, np.float64)
        self.L.orc_cp_set_slot_x(self.h, x, x.size // 3)

    def prune(self):
        """pruneNeighborCPU on the current lists; returns the entries removed."""
        return self.L.orc_cp_prune(self.h)

    def run(self, half, ntimes, reneigh, dt=0.005, dtforce=0.0025, mass=1.0, eps=1.0, sig6=1.0,
            cutf=6.25, prune_every=0):
        ks = np.zeros(ntimes + 1)
        rc = self.L.orc_cp_run_prune(self.h, int(half), eps, sig6, cutf, dt, dtforce, ntimes,
                                     reneigh, prune_every, mass, ks)
        assert rc == 0, rc
        return ks
