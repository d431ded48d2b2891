"""FP64 numpy restatement of the learned three-body surrogate (test infrastructure only).

The surrogate has no counterpart in KoMD: the nearest reference method is the
mean-field ATM2B scaling (potentials/ATM2B.cpp:9-20, restated in `atm2b_scale`
for the comparison tests).  What is pinned here is the mathematics the GPU
kernels must reproduce:

  descriptors  32 radial Gaussians and 16 angular invariants per site, within rc
               (rows of `INVARIANTS`), built from moment tensors
               T_a^L = sum_j psi_a(r_ij) u_ij^(x L); the invariant (a, b, L) is the
               full contraction <T_a^L, T_b^L> = sum_{j,k} psi_a(r_ij) psi_b(r_ik) cos^L
               (checked against that literal double sum in the tests)
  network      x = (D - shift) * scale; E_i = w3 . silu(W2 silu(W1 x + b1) + b2) + b3
  forces       F = -dE/dr with E = sum_i E_i (checked by central differences)
  training     L = wE/N sum_i (E_i - E*_i)^2 + wF/(3N) sum_i |F_i - F*_i|^2 and its
               exact gradient with respect to W1, b1, W2, b2, w3, b3 (checked by
               central differences of L)

Parameter vector layout (float, 7457 entries): W1[64][48], b1[64], W2[64][64],
b2[64], w3[64], b3[1], shift[48], scale[48].
"""
from __future__ import annotations

import numpy as np

RC = 2.5
NR = 32
R_MU0 = 0.85
R_DMU = 1.6 / 31.0
R_W = R_DMU
NA = 4
A_MU = (1.0, 1.35, 1.75, 2.15)
A_W = 0.35
A_POW = 3
# angular invariants (a, b, L): <T_a^L, T_b^L>
INVARIANTS = ([(0, 0, 1), (0, 1, 1), (0, 2, 1), (0, 3, 1), (1, 1, 1), (1, 2, 1)] +
              [(0, 0, 2), (0, 1, 2), (0, 2, 2), (0, 3, 2), (1, 1, 2), (1, 2, 2)] +
              [(0, 0, 3), (1, 1, 3), (2, 2, 3), (3, 3, 3)])
NDESC = NR + len(INVARIANTS)
H1 = 64
H2 = 64
NPARAM = H1 * NDESC + H1 + H2 * H1 + H2 + H2 + 1 + NDESC + NDESC


# ---- parameters -----------------------------------------------------------
def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    out = {}
    for name, shape in (("W1", (H1, NDESC)), ("b1", (H1,)), ("W2", (H2, H1)), ("b2", (H2,)), ("w3", (H2,)),
                        ("b3", (1,)), ("shift", (NDESC,)), ("scale", (NDESC,))):
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    assert o == NPARAM
    return out


def pack(d):
    return np.concatenate([np.asarray(d[k], np.float64).ravel() for k in
                           ("W1", "b1", "W2", "b2", "w3", "b3", "shift", "scale")])


def init_params(seed=0, shift=None, scale=None, e_mean=0.0, e_std=1.0):
    """Glorot-uniform weights (seeded), zero biases, output layer scaled to the target spread."""
    rng = np.random.default_rng(seed)
    lim1 = np.sqrt(6.0 / (NDESC + H1))
    lim2 = np.sqrt(6.0 / (H1 + H2))
    d = {"W1": rng.uniform(-lim1, lim1, (H1, NDESC)), "b1": np.zeros(H1),
         "W2": rng.uniform(-lim2, lim2, (H2, H1)), "b2": np.zeros(H2),
         "w3": rng.uniform(-1, 1, H2) * e_std / np.sqrt(H2), "b3": np.array([e_mean]),
         "shift": np.zeros(NDESC) if shift is None else shift,
         "scale": np.ones(NDESC) if scale is None else scale}
    return pack(d)


# ---- neighbours -------------------------------------------------------------
def neighbour_pairs(r, box, rc=RC):
    """Directed pairs (i, j), j != i, with |d| < rc, d = r_j - r_i (minimum image).
    Brute force in blocks: for the small test systems only."""
    r = np.asarray(r, np.float64)
    n = r.shape[0]
    I, J, Dv = [], [], []
    for s in range(0, n, 512):
        d = r[None, :, :] - r[s:s + 512, None, :]
        d -= box * np.rint(d / box)
        r2 = (d * d).sum(-1)
        ii, jj = np.nonzero(r2 < rc * rc)
        keep = (ii + s) != jj
        I.append(ii[keep] + s)
        J.append(jj[keep])
        Dv.append(d[ii[keep], jj[keep]])
    return np.concatenate(I), np.concatenate(J), np.concatenate(Dv)


def neighbour_pairs_cells(r, box, rc=RC, centers=None):
    """The pairs of neighbour_pairs from a cell list (cells of edge >= rc, 27 neighbour cells):
    the same pairs, the same FP64 displacements and the same (i, j) order, in O(N) work.
    centers restricts i to the given sites.  Boxes under three cells per side fall back to
    the brute force."""
    r = np.asarray(r, np.float64)
    box = np.asarray(box, np.float64)
    n = r.shape[0]
    nc = np.floor(box / rc).astype(np.int64)
    if (nc < 3).any():
        i, j, d = neighbour_pairs(r, box, rc)
        if centers is None:
            return i, j, d
        keep = np.isin(i, np.asarray(centers))
        return i[keep], j[keep], d[keep]
    cs = box / nc
    rw = r - box * np.floor(r / box)
    ci = np.minimum((rw / cs).astype(np.int64), nc - 1)
    flat = ci[:, 0] + nc[0] * (ci[:, 1] + nc[1] * ci[:, 2])
    ncell = int(np.prod(nc))
    cnt = np.bincount(flat, minlength=ncell)
    order = np.argsort(flat, kind="stable")
    start = np.concatenate([[0], np.cumsum(cnt)])
    m = int(cnt.max())
    table = np.full((ncell, m), -1, np.int64)          # the sites of every cell, padded with -1
    slot = np.arange(n) - start[flat[order]]
    table[flat[order], slot] = order
    cen = np.arange(n, dtype=np.int64) if centers is None else np.unique(np.asarray(centers, np.int64))
    I, J, Dv = [], [], []
    for ox in (-1, 0, 1):
        for oy in (-1, 0, 1):
            for oz in (-1, 0, 1):
                c = (ci[cen] + np.array([ox, oy, oz])) % nc
                cand = table[c[:, 0] + nc[0] * (c[:, 1] + nc[1] * c[:, 2])]     # [centres, m]
                ii = np.repeat(cen, m)
                jj = cand.ravel()
                ok = jj >= 0
                ii, jj = ii[ok], jj[ok]
                d = r[jj] - r[ii]
                d -= box * np.rint(d / box)
                r2 = (d * d).sum(-1)
                keep = (r2 < rc * rc) & (ii != jj)
                I.append(ii[keep])
                J.append(jj[keep])
                Dv.append(d[keep])
    I, J, Dv = np.concatenate(I), np.concatenate(J), np.concatenate(Dv)
    srt = np.lexsort((J, I))
    return I[srt], J[srt], Dv[srt]


# ---- basis functions ----------------------------------------------------------
def _fc(r):
    return 0.5 * (np.cos(np.pi * r / RC) + 1.0), -0.5 * np.pi / RC * np.sin(np.pi * r / RC)


def radial_basis(r):
    """phi_k(r) and dphi_k/dr, shape [P, 32]."""
    fc, dfc = _fc(r)
    mu = R_MU0 + R_DMU * np.arange(NR)
    x = (r[:, None] - mu[None]) / R_W
    g = np.exp(-x * x)
    return g * fc[:, None], g * (dfc[:, None] - 2.0 * x / R_W * fc[:, None])


def angular_basis(r):
    """psi_a(r) = exp(-((r - mu_a)/w)^2) fc(r) / r^3 and its derivative, shape [P, 4]."""
    fc, dfc = _fc(r)
    mu = np.asarray(A_MU)
    x = (r[:, None] - mu[None]) / A_W
    g = np.exp(-x * x)
    rp = r[:, None] ** (-A_POW)
    val = g * fc[:, None] * rp
    der = g * rp * (dfc[:, None] - 2.0 * x / A_W * fc[:, None] - A_POW / r[:, None] * fc[:, None])
    return val, der


def _outer(u, L):
    if L == 1:
        return u
    if L == 2:
        return np.einsum("pa,pb->pab", u, u)
    return np.einsum("pa,pb,pc->pabc", u, u, u)


# ---- descriptors --------------------------------------------------------------
def moments(n, i, d):
    """T[L] of shape [n, NA, 3^L] (full Cartesian tensors, flattened)."""
    r = np.linalg.norm(d, axis=1)
    u = d / r[:, None]
    psi, _ = angular_basis(r)
    T = {}
    for L in (1, 2, 3):
        t = _outer(u, L).reshape(len(r), -1)
        TL = np.zeros((n, NA, t.shape[1]))
        np.add.at(TL, i, psi[:, :, None] * t[:, None, :])
        T[L] = TL
    return T


def descriptors(n, i, d):
    """Raw descriptors D [n, 48] and the moment tensors."""
    r = np.linalg.norm(d, axis=1)
    phi, _ = radial_basis(r)
    D = np.zeros((n, NDESC))
    np.add.at(D[:, :NR], i, phi)
    T = moments(n, i, d)
    for m, (a, b, L) in enumerate(INVARIANTS):
        D[:, NR + m] = (T[L][:, a] * T[L][:, b]).sum(-1)
    return D, T


def angular_literal(n, i, d):
    """sum_{j,k in N(i)} psi_a(r_ij) psi_b(r_ik) cos^L(theta_jik), by explicit double loop."""
    r = np.linalg.norm(d, axis=1)
    u = d / r[:, None]
    psi, _ = angular_basis(r)
    out = np.zeros((n, len(INVARIANTS)))
    for s in range(n):
        sel = np.nonzero(i == s)[0]
        c = u[sel] @ u[sel].T
        for m, (a, b, L) in enumerate(INVARIANTS):
            out[s, m] = psi[sel, a] @ (c ** L) @ psi[sel, b]
    return out


# ---- network -------------------------------------------------------------------
def silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s * (1.0 + z * (1.0 - s)), s * (1.0 - s) * (2.0 + z * (1.0 - 2.0 * s))


def mlp(p, D):
    """Per-site energy and every intermediate (dict)."""
    P = unpack(p)
    x = (D - P["shift"]) * P["scale"]
    z1 = x @ P["W1"].T + P["b1"]
    h1, d1, dd1 = silu(z1)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2, d2, dd2 = silu(z2)
    E = h2 @ P["w3"] + P["b3"][0]
    t2 = d2 * P["w3"]                 # dE/dz2
    q = t2 @ P["W2"]                  # dE/dh1
    delta1 = d1 * q                   # dE/dz1
    g = delta1 @ P["W1"]              # dE/dx
    G = g * P["scale"]                # dE/dD
    return dict(x=x, z1=z1, h1=h1, d1=d1, dd1=dd1, z2=z2, h2=h2, d2=d2, dd2=dd2, E=E, t2=t2, q=q, delta1=delta1,
                g=g, G=G, P=P)


# ---- forces ----------------------------------------------------------------------
def _state(G, T):
    """A[L] [n, NA, 3^L]: dE/dT_a^L = sum over invariants containing a of G * T_other (x2 on the diagonal)."""
    A = {L: np.zeros_like(T[L]) for L in (1, 2, 3)}
    for m, (a, b, L) in enumerate(INVARIANTS):
        g = G[:, NR + m][:, None]
        A[L][:, a] += g * T[L][:, b]
        A[L][:, b] += g * T[L][:, a]
    return A


def _pair_grad(Grad, A, sites, d):
    """c = dE_site/dd for each pair (site, neighbour displacement d): [P, 3]."""
    r = np.linalg.norm(d, axis=1)
    u = d / r[:, None]
    _, dphi = radial_basis(r)
    psi, dpsi = angular_basis(r)
    c = (Grad[sites] * dphi).sum(1)[:, None] * u
    for L in (1, 2, 3):
        AL = A[L][sites].reshape(len(r), NA, *([3] * L))
        if L == 1:
            V = AL
        elif L == 2:
            V = np.einsum("paxy,py->pax", AL, u)
        else:
            V = np.einsum("paxyz,py,pz->pax", AL, u, u)
        s = np.einsum("pax,px->pa", V, u)
        c += np.einsum("pa,pa,px->px", dpsi, s, u)
        c += L * np.einsum("pa,pax->px", psi / r[:, None], V - s[:, :, None] * u[:, None, :])
    return c


def energy_forces(p, n, i, j, d):
    """Total energy, per-site energies and forces -dE/dr, FP64."""
    D, T = descriptors(n, i, d)
    m = mlp(p, D)
    A = _state(m["G"], T)
    Grad = m["G"][:, :NR]
    c_own = _pair_grad(Grad, A, i, d)        # dE_i/dd_ij
    c_nbr = _pair_grad(Grad, A, j, -d)       # dE_j/dd_ji
    F = np.zeros((n, 3))
    np.add.at(F, i, c_own - c_nbr)
    return float(m["E"].sum()), m["E"], F


def energy_forces_subset(p, r, box, sites, rc=RC):
    """Raw descriptors, per-site energies and forces -dE/dr of the given sites only, FP64.
    E_i needs the neighbours of i; F_i = sum_j (dE_i/dd_ij - dE_j/dd_ji) also needs the
    states of i's neighbours j, so the pairs of every site within rc of the subset are
    built (a cell list), the network runs on those sites, and the pair gradients only on
    the subset's pairs.  Equals energy_forces restricted to the subset."""
    r = np.asarray(r, np.float64)
    n = r.shape[0]
    sites = np.unique(np.asarray(sites, np.int64))
    i0, j0, _ = neighbour_pairs_cells(r, box, rc, centers=sites)
    ring = np.unique(np.concatenate([sites, j0]))
    i, j, d = neighbour_pairs_cells(r, box, rc, centers=ring)
    D, T = descriptors(n, i, d)
    m = mlp(p, D)
    A = _state(m["G"], T)
    own = np.isin(i, sites)
    Grad = m["G"][:, :NR]
    c_own = _pair_grad(Grad, A, i[own], d[own])
    c_nbr = _pair_grad(Grad, A, j[own], -d[own])
    F = np.zeros((n, 3))
    np.add.at(F, i[own], c_own - c_nbr)
    return D[sites], m["E"][sites], F[sites]


# ---- training loss and its gradient ------------------------------------------------
def _jvp(T, n, i, j, d, delta):
    """dD_a along the displacement field delta: sum_j dD_a/dd_aj . (delta_j - delta_a)."""
    r = np.linalg.norm(d, axis=1)
    u = d / r[:, None]
    _, dphi = radial_basis(r)
    psi, dpsi = angular_basis(r)
    dv = delta[j] - delta[i]
    ud = (u * dv).sum(1)
    w = dv - ud[:, None] * u                   # transverse part of the displacement
    dD = np.zeros((n, NDESC))
    np.add.at(dD[:, :NR], i, dphi * ud[:, None])
    dT = {}
    for L in (1, 2, 3):
        # d(psi u^L) = psi' (u.dv) u^L + psi/r * sum over slots of u..w..u
        base = _outer(u, L).reshape(len(r), -1)
        if L == 1:
            sym = w
        elif L == 2:
            sym = (np.einsum("pa,pb->pab", w, u) + np.einsum("pa,pb->pab", u, w)).reshape(len(r), -1)
        else:
            sym = (np.einsum("pa,pb,pc->pabc", w, u, u) + np.einsum("pa,pb,pc->pabc", u, w, u) +
                   np.einsum("pa,pb,pc->pabc", u, u, w)).reshape(len(r), -1)
        t = (dpsi * ud[:, None])[:, :, None] * base[:, None, :] + (psi / r[:, None])[:, :, None] * sym[:, None, :]
        dTL = np.zeros_like(T[L])
        np.add.at(dTL, i, t)
        dT[L] = dTL
    for m, (a, b, L) in enumerate(INVARIANTS):
        dD[:, NR + m] = (dT[L][:, a] * T[L][:, b]).sum(-1) + (T[L][:, a] * dT[L][:, b]).sum(-1)
    return dD


def loss_and_grad(p, n, i, j, d, E_ref, F_ref, wE, wF):
    """L = wE/N sum (E_i - E*_i)^2 + wF/(3N) sum |F_i - F*_i|^2 and dL/dp (shift, scale entries 0)."""
    D, T = descriptors(n, i, d)
    m = mlp(p, D)
    A = _state(m["G"], T)
    c_own = _pair_grad(m["G"][:, :NR], A, i, d)
    c_nbr = _pair_grad(m["G"][:, :NR], A, j, -d)
    F = np.zeros((n, 3))
    np.add.at(F, i, c_own - c_nbr)
    dE = m["E"] - E_ref
    dF = F - F_ref
    L = wE / n * float((dE * dE).sum()) + wF / (3 * n) * float((dF * dF).sum())
    eps = 2.0 * wE / n * dE                    # dL/dE_a
    delta = 2.0 * wF / (3 * n) * dF            # dL/dF_i
    P = m["P"]
    w = -P["scale"] * _jvp(T, n, i, j, d, delta)   # dL/dg_a (tangent direction in x)
    x, h1, h2, d1, d2, dd1, dd2 = m["x"], m["h1"], m["h2"], m["d1"], m["d2"], m["dd1"], m["dd2"]
    # energy path
    g_w3 = eps @ h2
    g_b3 = eps.sum()
    dz2E = eps[:, None] * d2 * P["w3"]
    # tangent path S = sum_a w_a . g_a
    u = w @ P["W1"].T
    dh1 = d1 * u
    dz2 = dh1 @ P["W2"].T
    t2 = m["t2"]
    e2 = P["w3"] * dd2 * dz2
    g_w3 = g_w3 + (d2 * dz2).sum(0)
    s2 = dz2E + e2                              # signal at z2 (energy + tangent)
    g_W2 = s2.T @ h1 + t2.T @ dh1
    g_b2 = s2.sum(0)
    q = m["q"]
    back = s2 @ P["W2"]                         # W2^T s2
    s1 = d1 * back + dd1 * u * q                # signal at z1
    g_W1 = s1.T @ x + m["delta1"].T @ w
    g_b1 = s1.sum(0)
    grad = pack({"W1": g_W1, "b1": g_b1, "W2": g_W2, "b2": g_b2, "w3": g_w3, "b3": np.array([g_b3]),
                 "shift": np.zeros(NDESC), "scale": np.zeros(NDESC)})
    return L, grad, m["E"], F


# ---- comparator: ATM2B mean-field scaling ------------------------------------------------
def atm2b_scale(nu, rho, eps=1.0, sig=1.0):
    """LJ scaling factor 1 - 2 nu rho / (3 eps sig^6) of the reference's ATM2B potential
    (potentials/ATM2B.cpp:9-20); its three-body force estimate is (scale - 1) * F_LJ."""
    return 1.0 - 2.0 * nu * rho / (3.0 * eps * sig ** 6)


# ---- hybrid: exact ATM in a slab, surrogate outside, cos^2 blend -----------------------------
def lam_of(x, axis_len, center, half, blend):
    """lambda = 1 for |x - c| <= h (minimum image), cos^2 descent to 0 over blend, 0 beyond;
    returns lambda and d lambda / dx."""
    d = x - center
    d = d - axis_len * np.rint(d / axis_len)
    ad = np.abs(d)
    lam = np.where(ad <= half, 1.0, 0.0)
    dl = np.zeros_like(x)
    mid = (ad > half) & (ad < half + blend)
    t = 0.5 * np.pi * (ad[mid] - half) / blend
    lam[mid] = np.cos(t) ** 2
    dl[mid] = -2.0 * np.cos(t) * np.sin(t) * (0.5 * np.pi / blend) * np.sign(d[mid])
    return lam, dl


def atm_triplets(r, box, rc=RC):
    """Every triplet i < j < k with all three minimum-image distances below rc."""
    i, j, d = neighbour_pairs_cells(r, box, rc)
    nb = [[] for _ in range(len(r))]
    for a, b in zip(i, j):
        nb[a].append(b)
    out = []
    for a in range(len(r)):
        js = np.array(sorted(x for x in nb[a] if x > a), dtype=np.int64)
        if len(js) < 2:
            continue
        jj, kk = np.meshgrid(js, js, indexing="ij")
        sel = jj < kk
        jj, kk = jj[sel], kk[sel]
        dv = r[kk] - r[jj]
        dv -= box * np.rint(dv / box)
        ok = (dv * dv).sum(1) < rc * rc
        out.append(np.stack([np.full(ok.sum(), a), jj[ok], kk[ok]], 1))
    return np.concatenate(out) if out else np.zeros((0, 3), np.int64)


def atm_triplet_terms(r, box, trip, nu):
    """V per triplet and dV/dr of each member (physical ATM), FP64 closed form:
    V = nu (1 + 3 cos1 cos2 cos3) / (r_ij r_ik r_jk)^3."""
    ri, rj, rk = r[trip[:, 0]], r[trip[:, 1]], r[trip[:, 2]]
    a = rj - ri
    a -= box * np.rint(a / box)
    b = rk - ri
    b -= box * np.rint(b / box)
    c = b - a
    A, B, C = (a * a).sum(1), (b * b).sum(1), (c * c).sum(1)
    p1, p2, p3 = (a * b).sum(1), (a * c).sum(1), (b * c).sum(1)
    Dd = A * B * C
    P = -p1 * p2 * p3
    V = nu * Dd ** -2.5 * (Dd + 3 * P)

    # gradient of V(a, b) with c = b - a, by the chain rule on (A, B, C, p1, p2, p3)
    dV_dD = nu * (-1.5 * Dd ** -2.5 - 7.5 * P * Dd ** -3.5)
    dV_dP = 3 * nu * Dd ** -2.5
    dD_dA, dD_dB, dD_dC = B * C, A * C, A * B
    dP_dp1, dP_dp2, dP_dp3 = -p2 * p3, -p1 * p3, -p1 * p2
    gA, gB, gC = dV_dD * dD_dA, dV_dD * dD_dB, dV_dD * dD_dC
    g1, g2, g3 = dV_dP * dP_dp1, dV_dP * dP_dp2, dV_dP * dP_dp3
    # dV/da and dV/db (a, b independent; c = b - a)
    dVa = (2 * gA[:, None] * a + g1[:, None] * b + g2[:, None] * (c - a) - g3[:, None] * b
           - 2 * gC[:, None] * c)
    dVb = (2 * gB[:, None] * b + g1[:, None] * a + g2[:, None] * a + g3[:, None] * (c + b)
           + 2 * gC[:, None] * c)
    # r_i enters as -(a + b), r_j as a, r_k as b
    return V, -(dVa + dVb), dVa, dVb


def hybrid_energy_forces(p, r, box, nu, axis, center, half, blend):
    """E = sum_t w_t V_t + sum_i (1 - lambda_i) E_i^sur, w_t = mean lambda of the members;
    F = -dE/dr including the gradients of the weights."""
    r = np.asarray(r, np.float64)
    n = len(r)
    lam, dl = lam_of(r[:, axis], box[axis], center, half, blend)
    F = np.zeros((n, 3))
    # exact part
    trip = atm_triplets(r, box)
    trip = trip[lam[trip].sum(1) > 0]   # triplets of zero weight add neither energy nor force
    E = 0.0
    if len(trip):
        V, gi, gj, gk = atm_triplet_terms(r, box, trip, nu)
        w = lam[trip].mean(1)
        E += float((w * V).sum())
        for col, g in ((0, gi), (1, gj), (2, gk)):
            np.add.at(F, trip[:, col], -w[:, None] * g)
            np.add.at(F[:, axis], trip[:, col], -V * dl[trip[:, col]] / 3.0)
    # surrogate part: site energies weighted by 1 - lambda
    cw = 1.0 - lam
    i, j, d = neighbour_pairs_cells(r, box)
    D, T = descriptors(n, i, d)
    m = mlp(p, D)
    G = m["G"] * cw[:, None]
    A = _state(G, T)
    c_own = _pair_grad(G[:, :NR], A, i, d)
    c_nbr = _pair_grad(G[:, :NR], A, j, -d)
    np.add.at(F, i, c_own - c_nbr)
    F[:, axis] += m["E"] * dl
    E += float((cw * m["E"]).sum())
    return E, F
