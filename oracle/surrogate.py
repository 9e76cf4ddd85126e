"""CPU oracle of the learned EAM surrogate (numpy, FP64).  Test infrastructure only: the tests,
the package's smoke check and bench.py's CPU baseline use it as the checker; the product never
calls it.

MD-Bench has no learned potential, so nothing here restates a reference file.  The functions
define the model the GPU kernels compute, and every one is pinned by what the mathematics fixes:

  descriptors          the O(n) moment form of the angular terms against the O(n^2) triple sum
                       sum_jk w_j w_k P_l(cos theta_jk) (brute force on small neighbourhoods);
                       radial terms against a direct per-pair loop; rotation and permutation
                       invariance
  mlp_forward/_grad    the input gradient against central differences of the energy
  forces               against central differences of the total energy sum_i E_i (ghost images
                       moved with their owners), with and without the core pair term
  force_adjoint,
  weight_gradients     against central differences of the loss in every parameter block
  adam_step            the textbook update (Kingma and Ba, 2015, Algorithm 1)

Geometry conventions: atom i's neighbours j come from a CSR list over local + ghost indices;
r_ij = x_j - x_i, u = r_ij / |r_ij|; a ghost g >= nlocal is an image of local atom
owner[g - nlocal].  Units: Angstrom, eV.
"""
import numpy as np

# ---------------------------------------------------------------- model constants
RC = 4.95                      # descriptor cutoff (the EAM table's cutoff)
NR = 24                        # Gaussian radial descriptors
R_MU0, R_DMU = 1.9, 0.13       # centres 1.90 ... 4.89 A
R_ETA = 1.0 / (2.0 * 0.16 ** 2)
NA = 4                         # angular channels (each gives an l=1 and an l=2 descriptor)
A_MU0, A_DMU = 2.5, 0.6        # channel centres 2.5, 3.1, 3.7, 4.3 A
A_ETA = 1.0 / (2.0 * 0.35 ** 2)
ND = NR + 2 * NA               # 32 descriptors
H = 64                         # hidden width
NCOEF = 64                     # force coefficients per atom: 24 + 4 x (3 + 6), padded
SYM = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]  # storage order of a symmetric 3x3


def cutoff(r):
    """fc(r) = (1 - (r/rc)^2)^3 inside rc, and dfc/dr."""
    x = (r / RC) ** 2
    t = np.where(x < 1.0, 1.0 - x, 0.0)
    return t ** 3, -6.0 * r / RC ** 2 * t ** 2


def radial_basis(r):
    """G_n(r) = exp(-eta (r - mu_n)^2) fc(r) and dG_n/dr, shape (..., NR)."""
    r = np.asarray(r, np.float64)[..., None]
    mu = R_MU0 + R_DMU * np.arange(NR)
    g = np.exp(-R_ETA * (r - mu) ** 2)
    fc, dfc = cutoff(r)
    return g * fc, g * (-2.0 * R_ETA * (r - mu) * fc + dfc)


def angular_weight(r):
    """w_c(r) = exp(-eta_a (r - mu_c)^2) fc(r) and dw_c/dr, shape (..., NA)."""
    r = np.asarray(r, np.float64)[..., None]
    mu = A_MU0 + A_DMU * np.arange(NA)
    g = np.exp(-A_ETA * (r - mu) ** 2)
    fc, dfc = cutoff(r)
    return g * fc, g * (-2.0 * A_ETA * (r - mu) * fc + dfc)


def _pairs(x, nlocal, numneigh, neighbors):
    """Flatten the CSR list to (i, j, r_ij vector, r) for the pairs inside RC."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    i = np.repeat(np.arange(nlocal), numneigh)
    j = np.asarray(neighbors[:i.size], np.int64)
    d = x[j] - x[i]
    r = np.sqrt((d * d).sum(1))
    m = r < RC
    return i[m], j[m], d[m], r[m]


def descriptors(x, nlocal, numneigh, neighbors):
    """D (nlocal, 32): 24 radial G_n, then q1_c = |A_c|^2 and q2_c = 1.5 |T~_c|_F^2 for the four
    channels, with A_c = sum_j w_c u_j, T_c = sum_j w_c u_j u_j^T and T~_c = T_c - tr(T_c) I / 3
    (1.5 |T~|^2 = 1.5 |T|^2 - 0.5 tr(T)^2 without the cancellation).  Also returns the moments A
    (nlocal, 4, 3) and T~ (nlocal, 4, 3, 3) the force pass needs.
    q1_c = sum_jk w_j w_k cos(theta_jk), q2_c = sum_jk w_j w_k P2(cos(theta_jk))."""
    i, j, d, r = _pairs(x, nlocal, numneigh, neighbors)
    u = d / r[:, None]
    G, _ = radial_basis(r)
    W, _ = angular_weight(r)
    D = np.zeros((nlocal, ND))
    np.add.at(D[:, :NR], i, G)
    A = np.zeros((nlocal, NA, 3))
    np.add.at(A, i, W[:, :, None] * u[:, None, :])
    T = np.zeros((nlocal, NA, 3, 3))
    np.add.at(T, i, W[:, :, None, None] * (u[:, :, None] * u[:, None, :])[:, None])
    D[:, NR:NR + NA] = (A * A).sum(-1)
    tr = np.trace(T, axis1=2, axis2=3)
    T = T - tr[..., None, None] * np.eye(3) / 3.0
    D[:, NR + NA:] = 1.5 * (T * T).sum((-1, -2))
    return D, A, T


def angular_triple_sum(x, i, nbrs):
    """q1 and q2 of atom i by the O(n^2) definition sum_jk w_j w_k P_l(cos theta_jk)."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    q1 = np.zeros(NA)
    q2 = np.zeros(NA)
    for a in nbrs:
        da = x[a] - x[i]
        ra = np.linalg.norm(da)
        if ra >= RC:
            continue
        wa, _ = angular_weight(ra)
        for b in nbrs:
            db = x[b] - x[i]
            rb = np.linalg.norm(db)
            if rb >= RC:
                continue
            wb, _ = angular_weight(rb)
            c = float(da @ db) / (ra * rb)
            q1 += wa * wb * c
            q2 += wa * wb * (1.5 * c * c - 0.5)
    return q1, q2


# ---------------------------------------------------------------- network
class Params:
    """MLP 32 -> 64 -> 64 -> 1 with tanh, on normalised descriptors.
    E_i = s_E (w3 . h2 + b3) + e0,  h2 = tanh(W2 h1 + b2),  h1 = tanh(W1 Dn + b1),
    Dn = (D - d_mean) / d_std."""

    NAMES = ("W1", "b1", "W2", "b2", "w3", "b3")

    def __init__(self, seed=1, d_mean=None, d_std=None, e0=0.0, s_e=1.0):
        rng = np.random.default_rng(seed)
        self.W1 = rng.normal(0, 1 / np.sqrt(ND), (H, ND))
        self.b1 = np.zeros(H)
        self.W2 = rng.normal(0, 1 / np.sqrt(H), (H, H))
        self.b2 = np.zeros(H)
        self.w3 = rng.normal(0, 1 / np.sqrt(H), H)
        self.b3 = np.zeros(1)
        self.d_mean = np.zeros(ND) if d_mean is None else np.asarray(d_mean, np.float64)
        self.d_std = np.ones(ND) if d_std is None else np.asarray(d_std, np.float64)
        self.e0 = float(e0)
        self.s_e = float(s_e)
        # the core pair term (not trained; k_core = 0: none)
        self.r_core = 0.0
        self.k_core = 0.0

    def flat(self):
        return np.concatenate([getattr(self, n).reshape(-1) for n in self.NAMES])

    def set_flat(self, v):
        o = 0
        for n in self.NAMES:
            a = getattr(self, n)
            setattr(self, n, np.asarray(v[o:o + a.size], np.float64).reshape(a.shape))
            o += a.size

    def copy(self):
        p = Params.__new__(Params)
        for n in self.NAMES + ("d_mean", "d_std"):
            setattr(p, n, getattr(self, n).copy())
        p.e0, p.s_e = self.e0, self.s_e
        p.r_core, p.k_core = self.r_core, self.k_core
        return p


NPARAM = H * ND + H + H * H + H + H + 1  # 6337


def bf16(a):
    """Round to bfloat16 (round to nearest even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def mlp_forward(p, D, bf16_inputs=False):
    """Per-atom energies and the cache the gradients use.  With bf16_inputs the matrix products
    take bfloat16-rounded operands (the GPU's MFMA inputs) and accumulate exactly."""
    q = bf16 if bf16_inputs else (lambda a: a)
    Dn = (D - p.d_mean) / p.d_std
    z1 = q(Dn) @ q(p.W1).T + p.b1
    h1 = np.tanh(z1)
    z2 = q(h1) @ q(p.W2).T + p.b2
    h2 = np.tanh(z2)
    E = p.s_e * (h2 @ p.w3 + p.b3[0]) + p.e0
    return E, dict(Dn=Dn, h1=h1, h2=h2)


def mlp_input_grad(p, cache, bf16_inputs=False):
    """g = dE_i/dD_i (nlocal, 32), through the normalisation."""
    q = bf16 if bf16_inputs else (lambda a: a)
    d2 = (1.0 - cache["h2"] ** 2) * p.w3
    d1 = (1.0 - cache["h1"] ** 2) * (q(d2) @ q(p.W2))
    gn = p.s_e * (q(d1) @ q(p.W1))
    return gn / p.d_std


# ---------------------------------------------------------------- forces
def coefficients(g, A, T):
    """Per-atom force coefficients (nlocal, 64): [24 radial g_n][4 x 3: 2 g_q1 A_c]
    [4 x 6: 3 g_q2 T~_c, symmetric storage xx yy zz xy xz yz][4 unused, zero].  T is the traceless
    moment from descriptors(); g . dq2/dx_j = 3 g T~ : dT/dx_j."""
    n = g.shape[0]
    c = np.zeros((n, NCOEF))
    c[:, :NR] = g[:, :NR]
    c[:, NR:NR + 3 * NA] = (2.0 * g[:, NR:NR + NA, None] * A).reshape(n, -1)
    B = 3.0 * g[:, NR + NA:, None, None] * T
    c[:, NR + 3 * NA:NR + 9 * NA] = np.stack([B[:, :, a, b] for a, b in SYM], -1).reshape(n, -1)
    return c


def _unpack(c):
    n = c.shape[0]
    rad = c[:, :NR]
    a = c[:, NR:NR + 3 * NA].reshape(n, NA, 3)
    s = c[:, NR + 3 * NA:NR + 9 * NA].reshape(n, NA, 6)
    B = np.zeros((n, NA, 3, 3))
    for k, (p_, q_) in enumerate(SYM):
        B[:, :, p_, q_] = s[:, :, k]
        B[:, :, q_, p_] = s[:, :, k]
    b = c[:, NR + 9 * NA:]
    return rad, a, B, b


def pair_grad(c, d, r):
    """v = dE_i/dx_j for pairs with displacement d = x_j - x_i, from atom i's coefficients c
    (one row per pair).  Sum of the radial, l=1 and l=2 terms."""
    u = d / r[:, None]
    _, dG = radial_basis(r)
    W, dW = angular_weight(r)
    rad, a, B, b = _unpack(c)
    s = (rad * dG).sum(1)
    ua = np.einsum("pk,pck->pc", u, a)
    Bu = np.einsum("pcab,pb->pca", B, u)
    uBu = np.einsum("pca,pa->pc", Bu, u)
    s = s + (dW * (ua + uBu - b)).sum(1)
    wr = W / r[:, None]
    t = (wr[:, :, None] * (a - ua[:, :, None] * u[:, None, :]
                           + 2.0 * (Bu - uBu[:, :, None] * u[:, None, :]))).sum(1)
    return s[:, None] * u + t


def owner_of(nlocal, owner):
    """Row of the coefficient table for every local + ghost index."""
    return np.concatenate([np.arange(nlocal), np.asarray(owner, np.int64)])


def forces(coef, x, nlocal, owner, numneigh, neighbors):
    """F_k = sum_{j in N(k)} (dE_k/dx_j - dE_j/dx_k) over the full list."""
    i, j, d, r = _pairs(x, nlocal, numneigh, neighbors)
    row = owner_of(nlocal, owner)
    v_ij = pair_grad(coef[i], d, r)
    v_ji = pair_grad(coef[row[j]], -d, r)
    F = np.zeros((nlocal, 3))
    np.add.at(F, i, v_ij - v_ji)
    return F


def core_energy_forces(x, nlocal, numneigh, neighbors, r_core, k_core):
    """The core pair term U(r) = k (r_core - r)^3 for r < r_core, 0 beyond (C2 at r_core): per
    atom E_i = 1/2 sum_j U(r_ij) and F_i = sum_j U'(r_ij) u_ij over the full list."""
    i, j, d, r = _pairs(x, nlocal, numneigh, neighbors)
    m = (r < r_core) & (r > 0)
    i, d, r = i[m], d[m], r[m]
    t = r_core - r
    E = np.zeros(nlocal)
    np.add.at(E, i, 0.5 * k_core * t ** 3)
    F = np.zeros((nlocal, 3))
    np.add.at(F, i, (-3.0 * k_core * t * t / r)[:, None] * d)
    return E, F


def energy_and_forces(p, x, nlocal, owner, numneigh, neighbors, bf16_inputs=False):
    D, A, T = descriptors(x, nlocal, numneigh, neighbors)
    E, cache = mlp_forward(p, D, bf16_inputs)
    g = mlp_input_grad(p, cache, bf16_inputs)
    F = forces(coefficients(g, A, T), x, nlocal, owner, numneigh, neighbors)
    if getattr(p, "k_core", 0.0) > 0.0:
        Ec, Fc = core_energy_forces(x, nlocal, numneigh, neighbors, p.r_core, p.k_core)
        E, F = E + Ec, F + Fc
    return E, F


# ---------------------------------------------------------------- training
def descriptor_jvp(x, nlocal, numneigh, neighbors, A, T, e_pair):
    """u_i[d] = sum_j e_ij . dD_id/dx_j: the directional derivative of atom i's descriptors when
    each neighbour j moves by e_ij (one row per pair inside RC, in _pairs order)."""
    i, j, d, r = _pairs(x, nlocal, numneigh, neighbors)
    u = d / r[:, None]
    ue = (u * e_pair).sum(1)
    _, dG = radial_basis(r)
    W, dW = angular_weight(r)
    out = np.zeros((nlocal, ND))
    np.add.at(out[:, :NR], i, dG * ue[:, None])
    Ai, Ti = A[i], T[i]
    uA = np.einsum("pk,pck->pc", u, Ai)
    eA = np.einsum("pk,pck->pc", e_pair, Ai)
    wr = W / r[:, None]
    q1 = 2.0 * (dW * ue[:, None] * uA + wr * (eA - ue[:, None] * uA))
    np.add.at(out[:, NR:NR + NA], i, q1)
    Tu = np.einsum("pcab,pb->pca", Ti, u)
    uTu = np.einsum("pca,pa->pc", Tu, u)
    eTu = np.einsum("pca,pa->pc", Tu, e_pair)
    q2 = 3.0 * (dW * uTu * ue[:, None] + 2.0 * wr * (eTu - uTu * ue[:, None]))
    np.add.at(out[:, NR + NA:], i, q2)
    return out


def force_adjoint(x, nlocal, owner, numneigh, neighbors, A, T, dF):
    """dL/dg_i for L = sum_k |dF_k|^2 with dF = F - F_ref: 2 sum_j J_ij^T (dF_i - dF_owner(j))."""
    i, j, _, _ = _pairs(x, nlocal, numneigh, neighbors)
    row = owner_of(nlocal, owner)
    e = dF[i] - dF[row[j]]
    return 2.0 * descriptor_jvp(x, nlocal, numneigh, neighbors, A, T, e)


def weight_gradients(p, cache, cE, ugrad):
    """d/dtheta of sum_i [cE_i E_i + ugrad_i . g_i] (cE_i = dL/dE_i, ugrad_i = dL/dg_i), by
    forward-mode along ugrad through the input gradient, then reverse mode."""
    Dn, h1, h2 = cache["Dn"], cache["h1"], cache["h2"]
    s = p.s_e
    ut = ugrad / p.d_std
    zd1 = ut @ p.W1.T
    t1 = 1.0 - h1 ** 2
    hd1 = t1 * zd1
    zd2 = hd1 @ p.W2.T
    t2 = 1.0 - h2 ** 2
    hd2 = t2 * zd2
    a2 = s * t2 * p.w3
    b2 = -2.0 * s * p.w3 * zd2 * h2 * t2
    g1 = a2 @ p.W2
    a1 = t1 * g1
    b1 = -2.0 * h1 * t1 * zd1 * g1 + t1 * (b2 @ p.W2)
    c = cE[:, None]
    out = Params.__new__(Params)
    out.W1 = (b1 + c * a1).T @ Dn + a1.T @ ut
    out.b1 = (b1 + c * a1).sum(0)
    out.W2 = (b2 + c * a2).T @ h1 + a2.T @ hd1
    out.b2 = (b2 + c * a2).sum(0)
    out.w3 = s * (hd2 + c * h2).sum(0)
    out.b3 = np.array([s * cE.sum()])
    return out


def loss(p, frames, wE, wF, bf16_inputs=False):
    """wE mean (E_i - E_ref)^2 + wF mean |F_k - F_ref|^2 over the frames' atoms (the components
    of F counted as one term each: mean over 3N)."""
    se, sf, n = 0.0, 0.0, 0
    for fr in frames:
        E, F = energy_and_forces(p, fr["x"], fr["nlocal"], fr["owner"], fr["numneigh"],
                                 fr["neighbors"], bf16_inputs)
        se += ((E - fr["e_ref"]) ** 2).sum()
        sf += ((F - fr["f_ref"].reshape(-1, 3)) ** 2).sum()
        n += fr["nlocal"]
    return wE * se / n + wF * sf / (3 * n)


def loss_gradient(p, frames, wE, wF):
    """Analytic gradient of loss() (FP64, no bf16 rounding), as a Params of the same shapes."""
    acc = None
    n = sum(fr["nlocal"] for fr in frames)
    for fr in frames:
        x, nl, ow, nn, nb = fr["x"], fr["nlocal"], fr["owner"], fr["numneigh"], fr["neighbors"]
        D, A, T = descriptors(x, nl, nn, nb)
        E, cache = mlp_forward(p, D)
        g = mlp_input_grad(p, cache)
        F = forces(coefficients(g, A, T), x, nl, ow, nn, nb)
        if getattr(p, "k_core", 0.0) > 0.0:  # the core: constant in the parameters
            Ec, Fc = core_energy_forces(x, nl, nn, nb, p.r_core, p.k_core)
            E, F = E + Ec, F + Fc
        dF = F - fr["f_ref"].reshape(-1, 3)
        ug = (wF / (3 * n)) * force_adjoint(x, nl, ow, nn, nb, A, T, dF)
        cE = (2.0 * wE / n) * (E - fr["e_ref"])
        gr = weight_gradients(p, cache, cE, ug)
        if acc is None:
            acc = gr
        else:
            for k in Params.NAMES:
                setattr(acc, k, getattr(acc, k) + getattr(gr, k))
    return acc


def adam_step(theta, grad, m, v, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba, 2015, Algorithm 1); t counts from 1."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
