"""The learned EAM surrogate's HIP kernels against the CPU oracle (oracle/surrogate.py), on Cu
snapshots the library's own plain-EAM path produces.

Tolerances: descriptors, moments and the force adjoint are FP32 sums (1e-5 relative); the network
runs on bfloat16 matrix cores, so the energies, coefficients and forces are compared with the
oracle evaluated on the same bfloat16-rounded operands (2e-3 relative RMS) and, loosely, with the
exact FP64 model; the weight gradient is FP32 and exact in form (1e-4)."""
import numpy as np
import pytest

from conftest import product, surrogate_module
from oracle import surrogate as so

pytestmark = pytest.mark.gpu


def _eam_frame(nx=4, temp=1600.0, steps=30):
    P = product()
    es = surrogate_module()
    sim = P.VLSim(nx=nx, ny=nx, nz=nx, force_field=1, lattice_a=3.615, temp=temp,
                  reneigh_every=10, ntimes=0)
    try:
        sim.run(steps)
        e = sim.eam_atom_energy()
        st = sim.state()
    finally:
        sim.close()
    return es._compact(st, e)


def _params_for(fr, seed=1):
    D, _, _ = so.descriptors(fr["x"], fr["nlocal"], fr["numneigh"], fr["neighbors"])
    p = so.Params(seed=seed, d_mean=D.mean(0), d_std=D.std(0) + 1e-3, e0=-3.5, s_e=0.3)
    rng = np.random.default_rng(seed + 10)
    p.b1 = rng.normal(0, 0.3, so.H)
    p.b2 = rng.normal(0, 0.3, so.H)
    p.b3 = rng.normal(0, 0.3, 1)
    # the device stores float32 weights: give the oracle the same values
    for k in so.Params.NAMES + ("d_mean", "d_std"):
        setattr(p, k, getattr(p, k).astype(np.float32).astype(np.float64))
    return p


def _model(es, p):
    return es.Model(p.flat().astype(np.float32), p.d_mean, p.d_std, p.e0, p.s_e, p.r_core,
                    p.k_core)


def _squeeze(frame, r_pair=1.7):
    """The frame with atom 0's nearest neighbour moved along the bond to r_pair (ghost images
    follow their owners)."""
    fr = dict(frame)
    nl = fr["nlocal"]
    x = fr["x"].reshape(-1, 3).copy()
    own = np.concatenate([np.arange(nl), fr["owner"]])
    shift = x - x[own]
    off = np.concatenate([[0], np.cumsum(fr["numneigh"])])
    j = fr["neighbors"][off[0]:off[1]]
    d = np.linalg.norm(x[j] - x[0], axis=1)
    k = own[j[np.argmin(d)]]
    u = (x[j[np.argmin(d)]] - x[0]) / d.min()
    x[k] -= (d.min() - r_pair) * u
    x = x[own] + shift
    fr["x"] = x.reshape(-1)
    return fr


def _lane_blocks(c):
    """The oracle's coefficient rows [24 radial][4 x 3 a][4 x 6 M][4 zero] in the device's lane
    blocks: block q = [radial 6q .. 6q+5, a_q, M_q, 0]."""
    out = np.zeros_like(c)
    for q in range(4):
        out[:, 16 * q:16 * q + 6] = c[:, 6 * q:6 * q + 6]
        out[:, 16 * q + 6:16 * q + 9] = c[:, 24 + 3 * q:24 + 3 * q + 3]
        out[:, 16 * q + 9:16 * q + 15] = c[:, 36 + 6 * q:36 + 6 * q + 6]
    return out


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_mfma_operand_layout_probe():
    """v_mfma_f32_16x16x32_bf16 as the network kernel uses it: C = A B on bfloat16 operands."""
    es = surrogate_module()
    rng = np.random.default_rng(0)
    A = rng.normal(size=(16, 32)).astype(np.float32)
    B = rng.normal(size=(32, 16)).astype(np.float32)
    Cg = np.zeros((16, 16), np.float32)
    assert es.lib().surMfmaProbeGFX950(es._p(A), es._p(B), es._p(Cg)) == 0
    ref = so.bf16(A) @ so.bf16(B)
    np.testing.assert_allclose(Cg, ref, rtol=1e-5, atol=1e-5)


@pytest.fixture(scope="module")
def frame():
    return _eam_frame()


def test_descriptors_and_moments_match_oracle(frame):
    es = surrogate_module()
    p = _params_for(frame)
    pk = es.PackedFrames([frame])
    ws = es.Workspace(pk.nlocal, pk.maxn)
    _model(es, p).upload()
    n = pk.nlocal
    assert es.lib().surDescriptorsGFX950(es.C.byref(pk.sys), 0, n, ws.dn.ptr, ws.mom.ptr, ws.cap,
                                         ws.pairs.ptr, ws.npairs.ptr) == 0
    # the compact pair list is the Verlet list cut at rc, in list order
    npairs = ws.npairs.get(n)
    np.testing.assert_array_equal(npairs, frame["numneigh"])
    pl = ws.pairs.get(n * ws.cap).reshape(ws.cap, n)
    csr = np.concatenate([pl[:npairs[a], a] for a in range(n)])
    np.testing.assert_array_equal(csr, frame["neighbors"])
    dn = ws.dn.get(n * so.ND).reshape(n, so.ND)
    mom = ws.mom.get(n * 36).reshape(n, 4, 9)  # per channel: A (3), traceless T (6)
    D, A, T = so.descriptors(frame["x"], n, frame["numneigh"], frame["neighbors"])
    Dn = (D - p.d_mean) / p.d_std
    # radial terms are O(1); compare each descriptor against its own scale
    scale = np.maximum(np.abs(D).max(0), 1e-6) / p.d_std
    assert np.max(np.abs(dn - Dn) / scale) < 2e-5
    np.testing.assert_allclose(mom[:, :, :3], A, atol=2e-5 * np.abs(A).max())
    Tg = mom[:, :, 3:]
    To = np.stack([T[:, :, a, b] for a, b in so.SYM], -1)
    np.testing.assert_allclose(Tg, To, atol=2e-5 * max(np.abs(To).max(), 1e-3))


def test_network_and_forces_match_oracle(frame):
    es = surrogate_module()
    p = _params_for(frame)
    pk = es.PackedFrames([frame])
    ws = es.Workspace(pk.nlocal, pk.maxn)
    _model(es, p).upload()
    n = pk.nlocal
    es.predict(pk.sys, 0, n, ws)
    e = ws.e.get(n)
    coef = ws.coef.get(n * 64).reshape(n, 64).astype(np.float64)
    f = ws.f.get(3 * n).reshape(n, 3)
    args = (frame["x"], n, frame["owner"], frame["numneigh"], frame["neighbors"])
    D, A, T = so.descriptors(frame["x"], n, frame["numneigh"], frame["neighbors"])
    Eb, cache = so.mlp_forward(p, D, bf16_inputs=True)
    gb = so.mlp_input_grad(p, cache, bf16_inputs=True)
    cb = so.coefficients(gb, A, T)
    Fb = so.forces(cb, *args)
    cb = _lane_blocks(cb)
    E64, F64 = so.energy_and_forces(p, *args)
    print("energy rel %.2e, coef rel %.2e, force rel (bf16 oracle) %.2e, (fp64) %.2e"
          % (_rel(e - p.e0, Eb - p.e0), _rel(coef, cb), _rel(f, Fb), _rel(f, F64)), flush=True)
    assert _rel(e - p.e0, Eb - p.e0) < 2e-3
    assert _rel(coef, cb) < 2e-3
    assert _rel(f, Fb) < 2e-3
    assert _rel(f, F64) < 5e-2
    assert np.abs(f).max() > 1e-2


def test_core_term_matches_oracle(frame):
    """The core pair term on a frame with a pair squeezed to 1.7 A: surCoreGFX950 (FP64) equals
    the oracle's, and the force call with the core set equals the network forces plus the core
    (bfloat16 oracle, 2e-3)."""
    es = surrogate_module()
    fr = _squeeze(frame)
    p = _params_for(frame)
    p.r_core, p.k_core = 2.0, 30.0
    pk = es.PackedFrames([fr])
    n = pk.nlocal
    L = es.lib()
    import ctypes as C
    ec, fc = es.Dev(n, np.float64), es.Dev(3 * n, np.float64)
    assert L.surCoreGFX950(C.byref(pk.sys), 0, n, 2.0, 30.0, ec.ptr, fc.ptr) == 0
    Ec, Fc = so.core_energy_forces(fr["x"], n, fr["numneigh"], fr["neighbors"], 2.0, 30.0)
    assert np.abs(Fc).max() > 1.0
    np.testing.assert_allclose(ec.get(n), Ec, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(fc.get(3 * n).reshape(n, 3), Fc, rtol=1e-12, atol=1e-12)
    ws = es.Workspace(n, pk.maxn)
    _model(es, p).upload()
    es.predict(pk.sys, 0, n, ws)
    f = ws.f.get(3 * n).reshape(n, 3)
    args = (fr["x"], n, fr["owner"], fr["numneigh"], fr["neighbors"])
    D, A, T = so.descriptors(fr["x"], n, fr["numneigh"], fr["neighbors"])
    _, cache = so.mlp_forward(p, D, bf16_inputs=True)
    Fb = so.forces(so.coefficients(so.mlp_input_grad(p, cache, bf16_inputs=True), A, T), *args)
    assert _rel(f, Fb + Fc) < 2e-3
    _model(es, p).upload(core=False)
    es.predict(pk.sys, 0, n, ws)
    assert _rel(ws.f.get(3 * n).reshape(n, 3), Fb) < 2e-3


def test_training_gradients_match_oracle(frame):
    """One step's adjoint and weight gradient against the oracle on the same inputs."""
    es = surrogate_module()
    L = es.lib()
    p = _params_for(frame)
    pk = es.PackedFrames([frame])
    n = pk.nlocal
    ws = es.Workspace(n, pk.maxn)
    _model(es, p).upload()
    es.predict(pk.sys, 0, n, ws)
    wE, wF = 1.0, 10.0
    sums = np.zeros(2)
    assert L.surLossTermsGFX950(n, ws.e.ptr, ws.f.ptr, pk.e_ref.ptr, pk.f_ref.ptr, wE, n,
                                ws.cE.ptr, ws.dF.ptr, es._p(sums)) == 0
    e = ws.e.get(n)
    f = ws.f.get(3 * n)
    cE = ws.cE.get(n).astype(np.float64)
    dF = ws.dF.get(3 * n).astype(np.float64).reshape(n, 3)
    np.testing.assert_allclose(cE, 2 * wE / n * (e - frame["e_ref"]), rtol=1e-5,
                               atol=1e-7 * np.abs(cE).max())
    np.testing.assert_allclose(dF.reshape(-1), f - frame["f_ref"], rtol=1e-6, atol=1e-6)
    assert abs(sums[1] - ((f - frame["f_ref"]) ** 2).sum()) < 1e-9 * sums[1]
    assert L.surAdjointGFX950(es.C.byref(pk.sys), 0, n, ws.pairs.ptr, ws.npairs.ptr, ws.mom.ptr,
                              ws.dF.ptr, wF / (3.0 * n), ws.U.ptr) == 0
    U = ws.U.get(n * 32).reshape(n, 32).astype(np.float64)
    D, A, T = so.descriptors(frame["x"], n, frame["numneigh"], frame["neighbors"])
    Uo = wF / (3.0 * n) * so.force_adjoint(frame["x"], n, frame["owner"], frame["numneigh"],
                                           frame["neighbors"], A, T, dF)
    print("adjoint rel %.2e" % _rel(U, Uo), flush=True)
    assert _rel(U, Uo) < 1e-4
    grad = es.Dev(so.NPARAM, np.float32)
    grad.zero()
    assert L.surWeightGradGFX950(n, ws.dn.ptr, ws.U.ptr, ws.cE.ptr, grad.ptr) == 0
    g = grad.get()
    dn = ws.dn.get(n * 32).reshape(n, 32).astype(np.float64)
    _, cache = so.mlp_forward(p, dn * p.d_std + p.d_mean)
    go = so.weight_gradients(p, cache, cE, U)
    off = 0
    for name in so.Params.NAMES:
        a = getattr(go, name).reshape(-1)
        r = _rel(g[off:off + a.size].astype(np.float64), a)
        print("grad %s rel %.2e" % (name, r), flush=True)
        assert r < 1e-4, name
        off += a.size


def test_adam_step_matches_oracle():
    es = surrogate_module()
    rng = np.random.default_rng(4)
    th = rng.normal(size=so.NPARAM).astype(np.float32)
    gr = rng.normal(size=so.NPARAM).astype(np.float32)
    m = (0.1 * rng.normal(size=so.NPARAM)).astype(np.float32)
    v = np.abs(0.1 * rng.normal(size=so.NPARAM)).astype(np.float32)
    dt, dg, dm, dv = (es.Dev.of(a, np.float32) for a in (th, gr, m, v))
    assert es.lib().surAdamGFX950(dt.ptr, dg.ptr, dm.ptr, dv.ptr, 7, 1e-3, 0.9, 0.999,
                                  1e-8) == 0
    # the kernel's constants are the float32 values of 0.9, 0.999 and 1e-3
    f32 = lambda z: float(np.float32(z))  # noqa: E731
    ref, mr, vr = so.adam_step(th.astype(np.float64), gr.astype(np.float64),
                               m.astype(np.float64), v.astype(np.float64), 7, lr=f32(1e-3),
                               beta1=f32(0.9), beta2=f32(0.999), eps=f32(1e-8))
    np.testing.assert_allclose(dm.get(), mr, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(dv.get(), vr, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(dt.get(), ref, rtol=1e-6, atol=1e-6)


@pytest.fixture(scope="module")
def small_training():
    """A short run: 4 temperatures x 10 snapshots of 6^3-cell boxes, 600 epochs (about 2 s)."""
    es = surrogate_module()
    model, info, data = es.train_surrogate(temps=[400.0, 900.0, 1400.0, 1900.0], per_temp=10,
                                           every=25, nx=6, epochs=600, batch_frames=2, wE=1.0,
                                           wF=1.0, lr0=3e-3, log=lambda s: print(s, flush=True))
    return model, info, data


def test_training_lowers_force_error_and_beats_linear_descriptor_fit(small_training):
    """What the trained network gains: its validation force error against the same network
    untrained, and against the method without a network, a least-squares linear fit of the
    per-atom energies on the same 32 descriptors (forces from the oracle)."""
    es = surrogate_module()
    model, info, (train, valid, frames, order, nval) = small_training
    vf = info["valid_force_rmse"]
    print("valid force RMSE trained %.1f meV/A, untrained %.1f meV/A"
          % (1e3 * vf, 1e3 * info["untrained_valid_force_rmse"]), flush=True)
    assert vf < 0.5 * info["untrained_valid_force_rmse"]
    # linear fit E_i = w . D_i + b on the training frames (oracle descriptors)
    Ds, Es = [], []
    for k in order[nval:]:
        fr = frames[k]
        D, _, _ = so.descriptors(fr["x"], fr["nlocal"], fr["numneigh"], fr["neighbors"])
        Ds.append(D)
        Es.append(fr["e_ref"])
    X = np.concatenate(Ds)
    X1 = np.hstack([X, np.ones((X.shape[0], 1))])
    w = np.linalg.lstsq(X1, np.concatenate(Es), rcond=None)[0]
    se, n = 0.0, 0
    for k in order[:nval]:
        fr = frames[k]
        nl = fr["nlocal"]
        _, A, T = so.descriptors(fr["x"], nl, fr["numneigh"], fr["neighbors"])
        g = np.tile(w[:so.ND], (nl, 1))
        F = so.forces(so.coefficients(g, A, T), fr["x"], nl, fr["owner"], fr["numneigh"],
                      fr["neighbors"])
        se += ((F.reshape(-1) - fr["f_ref"]) ** 2).sum()
        n += 3 * nl
    lin = np.sqrt(se / n)
    print("linear descriptor fit: valid force RMSE %.1f meV/A" % (1e3 * lin), flush=True)
    assert vf < lin


def test_drop_in_md_with_judge_and_fallback(small_training):
    """The surrogate as the EAM run's force call: judged every 10 steps; with an impossible
    tolerance the first judge falls back to plain EAM and the run goes on with EAM forces."""
    es = surrogate_module()
    P = product()
    model = small_training[0]
    sim = P.VLSim(nx=6, ny=6, nz=6, force_field=1, lattice_a=3.615, temp=1100.0,
                  reneigh_every=10, ntimes=0)
    try:
        es.use_in(sim, model, judge_every=10, tol=10.0)
        sim.run(40)
        st = es.judge_stats(sim)
        print("judge", st, flush=True)
        assert st["judges"] == 4 and not st["fallback"]
        assert 0.0 < st["max_rmse"] < 1.0
        es.use_in(sim, model, judge_every=10, tol=1e-9)
        sim.run(20)
        st = es.judge_stats(sim)
        assert st["fallback"] and st["fallback_step"] == 50
        # after the fallback the forces are plain EAM forces of the current positions
        f_run = sim.state(lists=False)["f"]
        sim.compute_force()
        f_eam = sim.state(lists=False)["f"]
        np.testing.assert_allclose(f_run, f_eam, rtol=1e-12, atol=1e-12)
    finally:
        sim.close()
