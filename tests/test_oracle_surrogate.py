"""The learned potential's CPU oracle (oracle/surrogate.py), pinned by what the mathematics fixes:
brute-force angular sums, invariances, and central differences of every derivative it claims
(forces, input gradients, the force adjoint and the weight gradients of the training loss)."""
import numpy as np
import pytest

import oracle_bind as ob
from oracle import surrogate as so

A_CU = 3.615
CUTNEIGH = 5.95


def fcc(nc, a=A_CU):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.stack(np.meshgrid(*[np.arange(nc)] * 3, indexing="ij"), -1).reshape(-1, 3)
    return ((cells[:, None, :] + base[None]) * a).reshape(-1, 3)


def periodic_frame(nc=3, disp=0.12, seed=0):
    """fcc Cu with Gaussian displacements, its ghost images and the full Verlet list (the
    oracle's restatement of MD-Bench's setupPbc / buildNeighbor)."""
    rng = np.random.default_rng(seed)
    L = nc * A_CU
    x = np.mod(fcc(nc) + rng.normal(0, disp, (4 * nc ** 3, 3)), L).reshape(-1)
    nl = x.size // 3
    prd = (L, L, L)
    bmap, pbc = ob.setup_pbc(nl, x, prd, CUTNEIGH)
    fr = dict(nlocal=nl, owner=bmap, pbc=pbc, prd=prd)
    fr["x"] = ghosts(fr, x)
    g = ob.vl_grid(nc, nc, nc, A_CU, CUTNEIGH)
    types = np.zeros(nl + bmap.size, np.int32)
    fr["numneigh"], fr["neighbors"] = ob.build_neighbor(g, nl, fr["x"], types, 1,
                                                        np.array([CUTNEIGH ** 2]), 0)
    return fr


def ghosts(fr, xloc):
    nl = fr["nlocal"]
    xall = np.zeros(3 * (nl + fr["owner"].size))
    xall[:3 * nl] = xloc[:3 * nl]
    return ob.update_pbc(nl, xall, fr["owner"], fr["pbc"], fr["prd"])


def cluster_frame(n=40, seed=3):
    """A free cluster (no ghosts) with an all-pairs list inside CUTNEIGH."""
    rng = np.random.default_rng(seed)
    x = fcc(3)[:n] + rng.normal(0, 0.1, (n, 3))
    d = np.linalg.norm(x[:, None] - x[None], axis=-1)
    nn, nb = [], []
    for i in range(n):
        js = [j for j in range(n) if j != i and d[i, j] < CUTNEIGH]
        nn.append(len(js))
        nb += js
    return dict(x=x.reshape(-1), nlocal=n, owner=np.zeros(0, np.int64),
                numneigh=np.array(nn, np.int32), neighbors=np.array(nb, np.int32))


def trained_like_params(fr, seed=1):
    """Random weights with the normalisation statistics of the frame (so activations are O(1))."""
    D, _, _ = so.descriptors(fr["x"], fr["nlocal"], fr["numneigh"], fr["neighbors"])
    p = so.Params(seed=seed, d_mean=D.mean(0), d_std=D.std(0) + 1e-3, e0=-3.5, s_e=0.3)
    rng = np.random.default_rng(seed + 10)
    p.b1 = rng.normal(0, 0.3, so.H)
    p.b2 = rng.normal(0, 0.3, so.H)
    p.b3 = rng.normal(0, 0.3, 1)
    return p


def test_radial_descriptors_match_direct_sum():
    fr = cluster_frame()
    D, _, _ = so.descriptors(fr["x"], fr["nlocal"], fr["numneigh"], fr["neighbors"])
    x = fr["x"].reshape(-1, 3)
    mu = so.R_MU0 + so.R_DMU * np.arange(so.NR)
    for i in [0, 7, 21]:
        ref = np.zeros(so.NR)
        for j in range(len(x)):
            r = np.linalg.norm(x[j] - x[i])
            if j != i and r < so.RC:
                ref += np.exp(-so.R_ETA * (r - mu) ** 2) * (1 - (r / so.RC) ** 2) ** 3
        np.testing.assert_allclose(D[i, :so.NR], ref, rtol=1e-12, atol=1e-14)


def test_angular_moments_equal_triple_sums():
    fr = cluster_frame()
    D, _, _ = so.descriptors(fr["x"], fr["nlocal"], fr["numneigh"], fr["neighbors"])
    off = np.concatenate([[0], np.cumsum(fr["numneigh"])])
    for i in [0, 5, 19, 33]:
        q1, q2 = so.angular_triple_sum(fr["x"], i, fr["neighbors"][off[i]:off[i + 1]])
        np.testing.assert_allclose(D[i, so.NR:so.NR + so.NA], q1, rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(D[i, so.NR + so.NA:], q2, rtol=1e-11, atol=1e-12)


def test_descriptors_rotation_and_permutation_invariant():
    fr = cluster_frame()
    D, _, _ = so.descriptors(fr["x"], fr["nlocal"], fr["numneigh"], fr["neighbors"])
    q, _ = np.linalg.qr(np.random.default_rng(7).normal(size=(3, 3)))
    xr = (fr["x"].reshape(-1, 3) @ q.T).reshape(-1)
    Dr, _, _ = so.descriptors(xr, fr["nlocal"], fr["numneigh"], fr["neighbors"])
    np.testing.assert_allclose(Dr, D, rtol=1e-10, atol=1e-12)
    # reversing every neighbour list changes nothing
    off = np.concatenate([[0], np.cumsum(fr["numneigh"])])
    nb = np.concatenate([fr["neighbors"][off[i]:off[i + 1]][::-1] for i in range(fr["nlocal"])])
    Dp, _, _ = so.descriptors(fr["x"], fr["nlocal"], fr["numneigh"], nb)
    np.testing.assert_allclose(Dp, D, rtol=1e-12, atol=1e-13)


def test_perfect_lattice_angular_terms_vanish_and_forces_are_zero():
    """fcc has cubic symmetry: the l=1 and l=2 moments vanish, and so do all forces."""
    fr = periodic_frame(disp=0.0)
    D, A, T = so.descriptors(fr["x"], fr["nlocal"], fr["numneigh"], fr["neighbors"])
    assert np.abs(A).max() < 1e-12
    assert np.abs(T).max() < 1e-12  # the traceless moment
    assert np.allclose(D, D[0], rtol=1e-12, atol=1e-12)
    p = trained_like_params(periodic_frame())
    _, F = so.energy_and_forces(p, fr["x"], fr["nlocal"], fr["owner"], fr["numneigh"],
                                fr["neighbors"])
    assert np.abs(F).max() < 1e-10


def test_input_gradient_matches_central_differences():
    fr = periodic_frame()
    p = trained_like_params(fr)
    D, _, _ = so.descriptors(fr["x"], fr["nlocal"], fr["numneigh"], fr["neighbors"])
    _, cache = so.mlp_forward(p, D)
    g = so.mlp_input_grad(p, cache)
    for d in [0, 5, 23, 24, 27, 28, 31]:
        h = 1e-6 * max(1.0, p.d_std[d])
        Dp, Dm = D.copy(), D.copy()
        Dp[:, d] += h
        Dm[:, d] -= h
        fd = (so.mlp_forward(p, Dp)[0] - so.mlp_forward(p, Dm)[0]) / (2 * h)
        np.testing.assert_allclose(g[:, d], fd, rtol=1e-6, atol=1e-8 * np.abs(fd).max())


@pytest.mark.parametrize("kind", ["cluster", "periodic"])
def test_forces_are_minus_energy_gradient(kind):
    fr = cluster_frame() if kind == "cluster" else periodic_frame()
    p = trained_like_params(fr)
    nl = fr["nlocal"]
    E, F = so.energy_and_forces(p, fr["x"], nl, fr["owner"], fr["numneigh"], fr["neighbors"])
    assert np.abs(F).max() > 1e-2
    h = 1e-5
    for k in [0, 3, nl // 2, nl - 1]:
        for c in range(3):
            es = []
            for sgn in (1, -1):
                x = fr["x"].copy()
                x[3 * k + c] += sgn * h
                if kind == "periodic":
                    x = ghosts(fr, x)
                es.append(so.energy_and_forces(p, x, nl, fr["owner"], fr["numneigh"],
                                               fr["neighbors"])[0].sum())
            fd = -(es[0] - es[1]) / (2 * h)
            assert abs(fd - F[k, c]) <= 1e-6 * max(1.0, np.abs(F).max()), (k, c, fd, F[k, c])


def squeezed_frame(seed=5, r_pair=1.7):
    """A periodic frame with one pair pushed to r_pair (inside a 2.0 A core)."""
    fr = periodic_frame(seed=seed)
    nl = fr["nlocal"]
    x = fr["x"].reshape(-1, 3).copy()
    off = np.concatenate([[0], np.cumsum(fr["numneigh"])])
    j = fr["neighbors"][off[0]:off[1]]
    d = np.linalg.norm(x[j] - x[0], axis=1)
    k = j[np.argmin(d)]
    k_own = k if k < nl else fr["owner"][k - nl]
    u = (x[k] - x[0]) / d.min()
    x[k_own] = x[k_own] - (d.min() - r_pair) * u  # move the neighbour's owner along the bond
    fr["x"] = ghosts(fr, x[:nl].reshape(-1))
    return fr


def test_core_term_is_c2_and_zero_beyond_rcore():
    """U(r) = k (rc - r)^3: value, slope and curvature vanish at rc and stay 0 beyond."""
    x = np.array([[0, 0, 0], [0, 0, 0.0]], np.float64)
    nn, nb = np.array([1, 1], np.int32), np.array([1, 0], np.int32)
    for r, e_exp in ((2.5, 0.0), (2.0, 0.0), (1.999, 0.5 * 30 * 1e-9), (1.5, 0.5 * 30 * 0.125)):
        x[1, 2] = r
        E, F = so.core_energy_forces(x.reshape(-1), 2, nn, nb, 2.0, 30.0)
        np.testing.assert_allclose(E, [e_exp, e_exp], rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(F[0], -F[1], rtol=0, atol=1e-15)
    x[1, 2] = 1.5
    _, F = so.core_energy_forces(x.reshape(-1), 2, nn, nb, 2.0, 30.0)
    assert F[0, 2] == pytest.approx(-3 * 30 * 0.25)  # pushed away from the neighbour at +z


def test_forces_with_core_are_minus_energy_gradient():
    """The network plus the core pair term on a frame with a pair at 1.7 A: central
    differences of the total energy."""
    fr = squeezed_frame()
    p = trained_like_params(periodic_frame())
    p.r_core, p.k_core = 2.0, 30.0
    nl = fr["nlocal"]
    E, F = so.energy_and_forces(p, fr["x"], nl, fr["owner"], fr["numneigh"], fr["neighbors"])
    Ec, Fc = so.core_energy_forces(fr["x"], nl, fr["numneigh"], fr["neighbors"], 2.0, 30.0)
    assert Ec.sum() > 0.1 and np.abs(Fc).max() > 1.0
    h = 1e-5
    for k in [0, int(np.argmax(np.abs(Fc).sum(1))), nl - 1]:
        for c in range(3):
            es = []
            for sgn in (1, -1):
                x = fr["x"].copy()
                x[3 * k + c] += sgn * h
                x = ghosts(fr, x)
                es.append(so.energy_and_forces(p, x, nl, fr["owner"], fr["numneigh"],
                                               fr["neighbors"])[0].sum())
            fd = -(es[0] - es[1]) / (2 * h)
            assert abs(fd - F[k, c]) <= 1e-6 * max(1.0, np.abs(F).max()), (k, c, fd, F[k, c])


def _frames_with_labels(n=2):
    frs = []
    for s in range(n):
        fr = periodic_frame(seed=11 + s)
        rng = np.random.default_rng(100 + s)
        fr["e_ref"] = -3.5 + rng.normal(0, 0.1, fr["nlocal"])
        fr["f_ref"] = rng.normal(0, 0.5, 3 * fr["nlocal"])
        frs.append(fr)
    return frs


def test_loss_gradient_matches_central_differences():
    """Energy + force loss: the analytic gradient (force adjoint pass, then forward-mode through
    the input gradient and reverse mode through the network) in every parameter block."""
    frs = _frames_with_labels()
    p = trained_like_params(frs[0])
    wE, wF = 1.0, 10.0
    gr = so.loss_gradient(p, frs, wE, wF)
    rng = np.random.default_rng(9)
    for name in so.Params.NAMES:
        a = getattr(p, name)
        for flat in rng.choice(a.size, size=min(4, a.size), replace=False):
            idx = np.unravel_index(flat, a.shape)
            h = 1e-6
            vals = []
            for sgn in (1, -1):
                q = p.copy()
                getattr(q, name)[idx] += sgn * h
                vals.append(so.loss(q, frs, wE, wF))
            fd = (vals[0] - vals[1]) / (2 * h)
            an = getattr(gr, name)[idx]
            assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-3), (name, idx, fd, an)


def test_force_adjoint_is_the_transpose_of_the_force_map():
    """<dF, J g> = <J^T dF, g> for the linear map from coefficients g to forces."""
    fr = periodic_frame(seed=4)
    nl = fr["nlocal"]
    D, A, T = so.descriptors(fr["x"], nl, fr["numneigh"], fr["neighbors"])
    rng = np.random.default_rng(2)
    g = rng.normal(size=(nl, so.ND))
    dF = rng.normal(size=(nl, 3))
    F = so.forces(so.coefficients(g, A, T), fr["x"], nl, fr["owner"], fr["numneigh"],
                  fr["neighbors"])
    u = so.force_adjoint(fr["x"], nl, fr["owner"], fr["numneigh"], fr["neighbors"], A, T, dF)
    lhs = 2.0 * (dF * F).sum()
    rhs = (u * g).sum()
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_adam_first_step_is_signed_learning_rate():
    """With m = v = 0, Adam's first step moves every parameter by lr * g / (|g| + eps)."""
    rng = np.random.default_rng(0)
    th = rng.normal(size=50)
    g = rng.normal(size=50)
    new, m, v = so.adam_step(th, g, np.zeros(50), np.zeros(50), 1, lr=0.01)
    np.testing.assert_allclose(th - new, 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1e-30], np.float64)
    r = so.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0             # exactly half an ulp (2^-7 spacing): to even
    assert r[2] == 1.0 + 2 ** -7   # above half: up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e38) / 3.0e38 < 2 ** -8
    assert abs(r[5] - 1e-30) / 1e-30 < 2 ** -8
