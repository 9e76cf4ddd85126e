"""Learned EAM surrogate: snapshot generation, GPU training and the drop-in force call.

Everything numerical runs in libmdgfx950's HIP kernels (include/mdgfx950_surrogate.h); this module
only moves arrays and sequences the calls:

  capture_frames   plain-EAM runs of small Cu boxes (the library's own EAM path), one snapshot
                   every `every` steps: positions with ghost images, the Verlet list, forces and
                   per-atom energies F(rho_i) + 1/2 sum phi as labels
  PackedFrames     many frames as one system on the device: locals of all frames, then all ghosts;
                   lists remapped to global indices (SoA, neighbors[k * nlocal + i])
  Trainer          Adam on wE mean (E - E_ref)^2 + wF mean |F - F_ref|^2: per batch of whole frames
                   descriptors -> bf16 MFMA network -> forces -> loss terms -> force adjoint ->
                   FP32 weight gradient -> Adam, all on the GPU
  use_in           hands the trained model to an EAM Verlet-list simulation
                   (mdbVLUseSurrogate: judged by plain EAM every 50 steps, fallback above 30 meV/A)
"""
import ctypes as C
import importlib.util
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

ND, NMOM, NCOEF, NPARAM = 32, 36, 64, 6337
H = 64
RC = 4.95
A_CU = 3.615
# Core pair term U(r) = K_CORE (R_CORE - r)^3 below R_CORE, outside what the network learns: the
# training snapshots hold no pair closer than about 1.93 A, and without it a pair that wanders
# below that range meets no repulsion (the network's extrapolation) and collapses.
R_CORE = 2.0   # A
K_CORE = 30.0  # eV / A^3


def _binding():
    if "mdbench_gfx950" in sys.modules:
        return sys.modules["mdbench_gfx950"]
    spec = importlib.util.spec_from_file_location("mdbench_gfx950",
                                                  os.path.join(HERE, "mdbench_gfx950.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mdbench_gfx950"] = mod
    spec.loader.exec_module(mod)
    return mod


md = _binding()


class SurSystem(C.Structure):
    _fields_ = [("nlocal", C.c_int), ("nghost", C.c_int), ("x", C.c_void_p),
                ("border_map", C.c_void_p), ("neighbors", C.c_void_p), ("numneigh", C.c_void_p)]


_bound = False


def lib():
    global _bound
    L = md.lib()
    if _bound:
        return L
    vp, i, f, d = C.c_void_p, C.c_int, C.c_float, C.c_double
    L.mdgDeviceAlloc.argtypes = [C.c_size_t]
    L.mdgDeviceAlloc.restype = vp
    L.mdgDeviceFree.argtypes = [vp]
    for fn in ("mdgCopyToDevice", "mdgCopyFromDevice"):
        getattr(L, fn).argtypes = [vp, vp, C.c_size_t]
        getattr(L, fn).restype = i
    L.mdgDeviceZero.argtypes = [vp, C.c_size_t]
    L.surSetModelGFX950.argtypes = [vp, vp, vp, f, f]
    L.surSetParamsDeviceGFX950.argtypes = [vp]
    L.surGetParamsGFX950.argtypes = [vp]
    sp = C.POINTER(SurSystem)
    L.surDescriptorsGFX950.argtypes = [sp, i, i, vp, vp, i, vp, vp]
    L.surEvaluateGFX950.argtypes = [i, vp, vp, vp, vp]
    L.surForcesGFX950.argtypes = [sp, i, i, vp, vp, vp, vp]
    L.surLossTermsGFX950.argtypes = [i, vp, vp, vp, vp, f, i, vp, vp, vp]
    L.surAdjointGFX950.argtypes = [sp, i, i, vp, vp, vp, vp, f, vp]
    L.surWeightGradGFX950.argtypes = [i, vp, vp, vp, vp]
    L.surAdamGFX950.argtypes = [vp, vp, vp, vp, i, f, f, f, f]
    L.surForceRmseGFX950.argtypes = [i, vp, vp]
    L.surForceRmseGFX950.restype = d
    L.surMfmaProbeGFX950.argtypes = [vp, vp, vp]
    L.surSetCoreGFX950.argtypes = [f, f]
    L.surCoreGFX950.argtypes = [sp, i, i, f, f, vp, vp]
    L.mdbVLUseSurrogate.argtypes = [vp, i, i, d]
    L.mdbVLSurrogateStats.argtypes = [vp, vp]
    for fn in ("surSetModelGFX950", "surSetParamsDeviceGFX950", "surGetParamsGFX950",
               "surDescriptorsGFX950", "surEvaluateGFX950", "surForcesGFX950",
               "surLossTermsGFX950", "surAdjointGFX950", "surWeightGradGFX950", "surAdamGFX950",
               "surMfmaProbeGFX950", "mdbVLUseSurrogate", "surSetCoreGFX950", "surCoreGFX950"):
        getattr(L, fn).restype = i
    _bound = True
    return L


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d): %s" % (what, rc, md.status()[1]))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Dev:
    """A device buffer (the library's hipMalloc)."""

    def __init__(self, count, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.ptr = lib().mdgDeviceAlloc(max(1, self.count) * self.dtype.itemsize)
        if not self.ptr:
            raise RuntimeError("device allocation of %d x %s failed: %s"
                               % (count, dtype, md.status()[1]))

    @classmethod
    def of(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype or a.dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size <= self.count
        _check(lib().mdgCopyToDevice(self.ptr, _p(a), a.nbytes), "upload")

    def get(self, count=None, offset=0):
        n = self.count - offset if count is None else count
        out = np.empty(n, self.dtype)
        _check(lib().mdgCopyFromDevice(_p(out), self.at(offset), out.nbytes), "download")
        return out

    def zero(self):
        _check(lib().mdgDeviceZero(self.ptr, ((self.count * self.dtype.itemsize + 3) // 4) * 4),
               "zero")

    def at(self, offset):
        return C.c_void_p(self.ptr + int(offset) * self.dtype.itemsize)

    def free(self):
        if self.ptr:
            lib().mdgDeviceFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


# ---------------------------------------------------------------- snapshots
def capture_frames(temps, per_temp, every=50, nx=10, log=None):
    """Plain-EAM runs (synthetic Cu tables, a = 3.615 A, dt = 1 fs, reneighbour every 10 steps)
    at each temperature from seeded Maxwell-Boltzmann velocities; a snapshot after every `every`
    steps.  Lists are cut to the descriptor
    cutoff."""
    frames = []
    for T in temps:
        sim = md.VLSim(nx=nx, ny=nx, nz=nx, force_field=1, lattice_a=A_CU, temp=float(T),
                       reneigh_every=10, ntimes=0, maxwell_seed=int(T) + 1)
        try:
            for _ in range(per_temp):
                sim.run(every)
                e = sim.eam_atom_energy()
                st = sim.state()
                frames.append(_compact(st, e))
        finally:
            sim.close()
        if log:
            log("captured %d frames at %.0f K" % (per_temp, T))
    return frames


def _compact(st, e):
    nl = st["nlocal"]
    x = st["x"]
    nn = st["numneigh"].astype(np.int64)
    nb = st["neighbors"][:nn.sum()].astype(np.int64)
    i = np.repeat(np.arange(nl), nn)
    xr = x.reshape(-1, 3)
    d = xr[nb] - xr[i]
    keep = (d * d).sum(1) < RC * RC
    nn2 = np.bincount(i[keep], minlength=nl).astype(np.int32)
    return dict(nlocal=nl, x=x.copy(), owner=st["border_map"].astype(np.int32).copy(),
                numneigh=nn2, neighbors=nb[keep].astype(np.int32), e_ref=e.copy(),
                f_ref=st["f"].copy())


class PackedFrames:
    """Frames as one device-resident system (locals of all frames, then all ghosts)."""

    def __init__(self, frames):
        nls = np.array([f["nlocal"] for f in frames], np.int64)
        ngs = np.array([f["owner"].size for f in frames], np.int64)
        self.nlocal = int(nls.sum())
        self.nghost = int(ngs.sum())
        self.loff = np.concatenate([[0], np.cumsum(nls)])
        goff = np.concatenate([[0], np.cumsum(ngs)])
        maxn = max(int(f["numneigh"].max()) for f in frames)
        N = self.nlocal
        x = np.empty(3 * (N + self.nghost))
        bmap = np.empty(max(self.nghost, 1), np.int32)
        nb = np.zeros(maxn * N, np.int32)
        nn = np.empty(N, np.int32)
        e = np.empty(N)
        f = np.empty(3 * N)
        for k, fr in enumerate(frames):
            nl, ng, L0, G0 = fr["nlocal"], fr["owner"].size, self.loff[k], goff[k]
            x[3 * L0:3 * (L0 + nl)] = fr["x"][:3 * nl]
            x[3 * (N + G0):3 * (N + G0 + ng)] = fr["x"][3 * nl:3 * (nl + ng)]
            bmap[G0:G0 + ng] = L0 + fr["owner"]
            cnt = fr["numneigh"].astype(np.int64)
            j = fr["neighbors"].astype(np.int64)
            jg = np.where(j < nl, L0 + j, N + G0 + (j - nl))
            ia = np.repeat(np.arange(nl), cnt)
            slot = np.arange(cnt.sum()) - np.repeat(np.concatenate([[0], np.cumsum(cnt)[:-1]]),
                                                    cnt)
            nb[slot * N + L0 + ia] = jg
            nn[L0:L0 + nl] = fr["numneigh"]
            e[L0:L0 + nl] = fr["e_ref"]
            f[3 * L0:3 * (L0 + nl)] = fr["f_ref"]
        self.maxn = maxn
        self.x = Dev.of(x, np.float64)
        self.bmap = Dev.of(bmap, np.int32)
        self.nb = Dev.of(nb, np.int32)
        self.nn = Dev.of(nn, np.int32)
        self.e_ref = Dev.of(e, np.float64)
        self.f_ref = Dev.of(f, np.float64)
        self.e_host, self.f_host = e, f
        self.sys = SurSystem(N, self.nghost, self.x.ptr, self.bmap.ptr, self.nb.ptr, self.nn.ptr)

    def subtract_core(self, r_core, k_core):
        """Labels minus the core pair term (computed on the device over the packed lists)."""
        N = self.nlocal
        ec, fc = Dev(N, np.float64), Dev(3 * N, np.float64)
        _check(lib().surCoreGFX950(C.byref(self.sys), 0, N, r_core, k_core, ec.ptr, fc.ptr),
               "surCoreGFX950")
        self.e_host = self.e_host - ec.get()
        self.f_host = self.f_host - fc.get()
        self.e_ref = Dev.of(self.e_host, np.float64)
        self.f_ref = Dev.of(self.f_host, np.float64)

    def frame_range(self, a, b):
        """Atom range [i0, i1) of frames a .. b-1."""
        return int(self.loff[a]), int(self.loff[b])

    @property
    def nframes(self):
        return len(self.loff) - 1


# ---------------------------------------------------------------- model
def init_theta(seed=1):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.normal(0, 1 / np.sqrt(ND), H * ND), np.zeros(H),
                           rng.normal(0, 1 / np.sqrt(H), H * H), np.zeros(H),
                           rng.normal(0, 1 / np.sqrt(H), H), np.zeros(1)]).astype(np.float32)


class Model:
    """Weights and normalisation of the network (host copies) and the core pair term."""

    def __init__(self, theta, d_mean, d_std, e0, s_e, r_core=R_CORE, k_core=K_CORE):
        self.theta = np.asarray(theta, np.float32)
        self.d_mean = np.asarray(d_mean, np.float32)
        self.d_std = np.asarray(d_std, np.float32)
        self.e0, self.s_e = float(e0), float(s_e)
        self.r_core, self.k_core = float(r_core), float(k_core)

    def upload(self, core=True):
        """Network and normalisation to the device; core=False leaves the core term out of the
        force call (training fits the network to labels with the core subtracted)."""
        _check(lib().surSetModelGFX950(_p(self.theta), _p(self.d_mean), _p(self.d_std),
                                       self.e0, self.s_e), "surSetModelGFX950")
        _check(lib().surSetCoreGFX950(self.r_core, self.k_core if core else 0.0),
               "surSetCoreGFX950")

    def save(self, path):
        np.savez(path, theta=self.theta, d_mean=self.d_mean, d_std=self.d_std,
                 e0=self.e0, s_e=self.s_e, r_core=self.r_core, k_core=self.k_core)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        core = (float(z["r_core"]), float(z["k_core"])) if "k_core" in z.files else (0.0, 0.0)
        return cls(z["theta"], z["d_mean"], z["d_std"], float(z["e0"]), float(z["s_e"]), *core)


class Workspace:
    """Per-batch device arrays for up to n atoms whose lists hold at most cap entries."""

    def __init__(self, n, cap):
        self.n = n
        self.cap = int(cap)
        self.dn = Dev(n * ND, np.float32)
        self.mom = Dev(n * NMOM, np.float32)
        self.coef = Dev(n * NCOEF, np.float16)
        self.pairs = Dev(n * self.cap, np.int32)
        self.npairs = Dev(n, np.int32)
        self.e = Dev(n, np.float32)
        self.f = Dev(3 * n, np.float64)
        self.cE = Dev(n, np.float32)
        self.dF = Dev(3 * n, np.float32)
        self.U = Dev(n * ND, np.float32)


def predict(sysobj, i0, n, ws):
    """Descriptors, network and forces of atoms i0 .. i0+n-1 into the workspace."""
    L = lib()
    s = C.byref(sysobj)
    _check(L.surDescriptorsGFX950(s, i0, n, ws.dn.ptr, ws.mom.ptr, ws.cap, ws.pairs.ptr,
                                  ws.npairs.ptr), "descriptors")
    _check(L.surEvaluateGFX950(n, ws.dn.ptr, ws.mom.ptr, ws.e.ptr, ws.coef.ptr), "evaluate")
    _check(L.surForcesGFX950(s, i0, n, ws.pairs.ptr, ws.npairs.ptr, ws.coef.ptr, ws.f.ptr),
           "forces")


def descriptor_stats(packed, frames_sample, ws):
    """Mean and standard deviation of the raw descriptors over a sample of frames."""
    ident = Model(init_theta(), np.zeros(ND), np.ones(ND), 0.0, 1.0)
    ident.upload()
    acc = []
    for k in frames_sample:
        i0, i1 = packed.frame_range(k, k + 1)
        _check(lib().surDescriptorsGFX950(C.byref(packed.sys), i0, i1 - i0, ws.dn.ptr,
                                          ws.mom.ptr, ws.cap, ws.pairs.ptr, ws.npairs.ptr),
               "descriptors")
        acc.append(ws.dn.get((i1 - i0) * ND).reshape(-1, ND))
    D = np.concatenate(acc).astype(np.float64)
    return D.mean(0), D.std(0) + 1e-6


class Trainer:
    def __init__(self, train, valid, model, batch_frames=16, wE=1.0, wF=1.0, log=None):
        self.train, self.valid = train, valid
        self.model = model
        self.batch_frames = batch_frames
        self.wE, self.wF = wE, wF
        self.log = log or (lambda s: None)
        nmax = max(self._max_batch_atoms(train), self._max_batch_atoms(valid))
        self.ws = Workspace(nmax, max(train.maxn, valid.maxn))
        self.theta = Dev.of(model.theta, np.float32)
        self.m = Dev(NPARAM, np.float32)
        self.v = Dev(NPARAM, np.float32)
        self.grad = Dev(NPARAM, np.float32)
        self.m.zero()
        self.v.zero()
        self.t = 0
        model.upload(core=False)

    def _max_batch_atoms(self, pk):
        b = self.batch_frames
        return max(pk.loff[min(k + b, pk.nframes)] - pk.loff[k] for k in range(0, pk.nframes, b))

    def step(self, i0, n, lr):
        L = lib()
        ws, pk = self.ws, self.train
        predict(pk.sys, i0, n, ws)
        sums = np.zeros(2)
        _check(L.surLossTermsGFX950(n, ws.e.ptr, ws.f.ptr, pk.e_ref.at(i0), pk.f_ref.at(3 * i0),
                                    self.wE, n, ws.cE.ptr, ws.dF.ptr, _p(sums)), "loss terms")
        _check(L.surAdjointGFX950(C.byref(pk.sys), i0, n, ws.pairs.ptr, ws.npairs.ptr, ws.mom.ptr,
                                  ws.dF.ptr, self.wF / (3.0 * n), ws.U.ptr), "adjoint")
        self.grad.zero()
        _check(L.surWeightGradGFX950(n, ws.dn.ptr, ws.U.ptr, ws.cE.ptr, self.grad.ptr), "wgrad")
        self.t += 1
        _check(L.surAdamGFX950(self.theta.ptr, self.grad.ptr, self.m.ptr, self.v.ptr, self.t, lr,
                               0.9, 0.999, 1e-8), "adam")
        _check(L.surSetParamsDeviceGFX950(self.theta.ptr), "set params")
        return sums[0] / n, sums[1] / (3 * n)

    def evaluate(self, pk):
        """Energy RMSE (eV/atom) and force RMSE per component (eV/A) over all frames of pk."""
        L = lib()
        se = sf = 0.0
        na = 0
        for a in range(0, pk.nframes, self.batch_frames):
            i0, i1 = pk.frame_range(a, min(a + self.batch_frames, pk.nframes))
            n = i1 - i0
            predict(pk.sys, i0, n, self.ws)
            sums = np.zeros(2)
            _check(L.surLossTermsGFX950(n, self.ws.e.ptr, self.ws.f.ptr, pk.e_ref.at(i0),
                                        pk.f_ref.at(3 * i0), 1.0, n, self.ws.cE.ptr,
                                        self.ws.dF.ptr, _p(sums)), "loss terms")
            se += sums[0]
            sf += sums[1]
            na += n
        return np.sqrt(se / na), np.sqrt(sf / (3 * na))

    def fit(self, epochs, lr0=2e-3, lr1=5e-5, seed=0):
        rng = np.random.default_rng(seed)
        nb = (self.train.nframes + self.batch_frames - 1) // self.batch_frames
        total = epochs * nb
        hist = []
        t0 = time.perf_counter()
        k = 0
        for ep in range(epochs):
            el = fl = 0.0
            for b in rng.permutation(nb):
                a = b * self.batch_frames
                i0, i1 = self.train.frame_range(a, min(a + self.batch_frames, self.train.nframes))
                lr = lr0 * (lr1 / lr0) ** (k / max(1, total - 1))
                e2, f2 = self.step(i0, i1 - i0, lr)
                el += e2
                fl += f2
                k += 1
            ve, vf = self.evaluate(self.valid)
            hist.append((ep, np.sqrt(el / nb), np.sqrt(fl / nb), ve, vf))
            if ep % max(1, epochs // 15) and ep != epochs - 1:
                continue
            self.log("epoch %3d  train E %.2f meV/atom F %.1f meV/A   valid E %.2f meV/atom "
                     "F %.1f meV/A  (%.1f s)" % (ep, 1e3 * hist[-1][1], 1e3 * hist[-1][2],
                                                  1e3 * ve, 1e3 * vf, time.perf_counter() - t0))
        self.model.theta = self.theta.get()
        return hist


def train_surrogate(temps=None, per_temp=100, every=50, nx=10, epochs=30, batch_frames=16,
                    valid_fraction=0.1, seed=0, wE=1.0, wF=1.0, lr0=2e-3, lr1=5e-5, log=None):
    """Snapshots from plain EAM, a 90/10 split, Adam on the GPU.  Returns (model, info)."""
    log = log or (lambda s: None)
    temps = np.linspace(300.0, 2000.0, 20) if temps is None else temps
    t0 = time.perf_counter()
    frames = capture_frames(temps, per_temp, every=every, nx=nx, log=log)
    t_cap = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(frames))
    nval = max(1, int(round(valid_fraction * len(frames))))
    valid = PackedFrames([frames[k] for k in order[:nval]])
    train = PackedFrames([frames[k] for k in order[nval:]])
    for pk in (train, valid):
        pk.subtract_core(R_CORE, K_CORE)
    ws0 = Workspace(int(max(f["nlocal"] for f in frames)), train.maxn)
    d_mean, d_std = descriptor_stats(train, range(0, train.nframes, max(1, train.nframes // 64)),
                                     ws0)
    del ws0
    e0, s_e = float(train.e_host.mean()), float(train.e_host.std() + 1e-6)
    model = Model(init_theta(seed + 1), d_mean, d_std, e0, s_e)
    tr = Trainer(train, valid, model, batch_frames=batch_frames, wE=wE, wF=wF, log=log)
    e_init, f_init = tr.evaluate(valid)
    t1 = time.perf_counter()
    hist = tr.fit(epochs, lr0=lr0, lr1=lr1, seed=seed)
    t_fit = time.perf_counter() - t1
    info = dict(frames=len(frames), train_frames=train.nframes, valid_frames=valid.nframes,
                train_atoms=train.nlocal, capture_s=t_cap, fit_s=t_fit, epochs=epochs,
                untrained_valid_force_rmse=f_init, untrained_valid_energy_rmse=e_init,
                valid_energy_rmse=hist[-1][3], valid_force_rmse=hist[-1][4], history=hist)
    return model, info, (train, valid, frames, order, nval)


def use_in(sim, model, judge_every=50, tol=0.030):
    """Switch an EAM VLSim to the surrogate (judged by plain EAM every judge_every steps)."""
    model.upload()
    _check(lib().mdbVLUseSurrogate(sim.h, 1, judge_every, tol), "mdbVLUseSurrogate")


def judge_stats(sim):
    out = np.zeros(5)
    lib().mdbVLSurrogateStats(sim.h, _p(out))
    return dict(judges=int(out[0]), last_rmse=out[1], max_rmse=out[2],
                fallback_step=int(out[3]), fallback=bool(out[4]))
