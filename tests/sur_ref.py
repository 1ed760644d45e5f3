"""FP64 numpy restatement of the Ewald surrogate (md-lab_amd/csrc/surrogate.hip),
test infrastructure only: descriptors, network, energies, analytic forces
and the training-loss gradient, used to check the HIP kernels (which run
the descriptors in FP32 and the network in BF16 with FP32 accumulation).

Per bead slice of N particles in a periodic cubic box, with u_il the unit
vector and r_il the distance from particle i to the minimum image of l,
fc(r) = (cos(pi r / r_c) + 1) / 2 below r_c:
  radial   G_f = sum_l exp(-eta_r (r_il - f dmu_r)^2) fc(r_il), f = 0..63,
           dmu_r = r_c / 64, eta_r = 1 / (2 (2 dmu_r)^2)
  angular  per channel c = 0..7: w_c(r) = exp(-eta_a (r - (c+1/2) dmu_a)^2) fc(r),
           dmu_a = r_c / 8, eta_a = 1 / (2 dmu_a^2);
           M0 = sum w, M1 = sum w u, M2 = sum w u u^T;
           features M0^2 + |M1|^2, M0^2 - |M1|^2, M0^2 + 2|M1|^2 + |M2|_F^2,
           M0^2 - 2|M1|^2 + |M2|_F^2
  network  e_i = (w3 . tanh(W2^T tanh(W1^T Dn_i + b1) + b2) + b3) ysd + ymu,
           Dn = (D - mean) * istd
"""
import numpy as np

NR, NCH, H = 64, 8, 128
ND = NR + 4 * NCH
NPAR = ND * H + H + H * H + H + H + 1


def split_params(par):
    par = np.asarray(par, dtype=np.float64)
    o = 0
    W1 = par[o:o + ND * H].reshape(ND, H); o += ND * H
    b1 = par[o:o + H]; o += H
    W2 = par[o:o + H * H].reshape(H, H); o += H * H
    b2 = par[o:o + H]; o += H
    w3 = par[o:o + H]; o += H
    b3 = par[o]
    return W1, b1, W2, b2, w3, b3


def geometry(rc):
    dmu_r = rc / NR
    dmu_a = rc / NCH
    return dict(rc=rc, dmu_r=dmu_r, eta_r=1.0 / (2 * (2 * dmu_r) ** 2), dmu_a=dmu_a, eta_a=1.0 / (2 * dmu_a ** 2))


def pairs(x, L, rc):
    """Separations r_il = x_l - x_i (minimum image) of all pairs within rc:
    (i, l, rvec, r)."""
    d = x[None, :, :] - x[:, None, :]
    d -= L * np.rint(d / L)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    i, l = np.nonzero(r < rc)
    return i, l, d[i, l], r[i, l]


def radial_basis(r, g):
    fc = 0.5 * (np.cos(np.pi * r / g["rc"]) + 1)
    dfc = -0.5 * np.pi / g["rc"] * np.sin(np.pi * r / g["rc"])
    mu = np.arange(NR) * g["dmu_r"]
    e = np.exp(-g["eta_r"] * (r[:, None] - mu) ** 2)
    gv = e * fc[:, None]
    dg = e * (dfc[:, None] - 2 * g["eta_r"] * (r[:, None] - mu) * fc[:, None])
    return gv, dg


def angular_basis(r, g):
    fc = 0.5 * (np.cos(np.pi * r / g["rc"]) + 1)
    dfc = -0.5 * np.pi / g["rc"] * np.sin(np.pi * r / g["rc"])
    mu = (np.arange(NCH) + 0.5) * g["dmu_a"]
    e = np.exp(-g["eta_a"] * (r[:, None] - mu) ** 2)
    return e * fc[:, None], e * (dfc[:, None] - 2 * g["eta_a"] * (r[:, None] - mu) * fc[:, None])


def descriptors(x, L, rc):
    """Raw descriptors [N][96] and moments (M0 [N][8], M1 [N][8][3], M2 [N][8][3][3])."""
    g = geometry(rc)
    N = len(x)
    i, l, rv, r = pairs(x, L, rc)
    u = rv / r[:, None]
    gv, _ = radial_basis(r, g)
    w, _ = angular_basis(r, g)
    G = np.zeros((N, NR))
    np.add.at(G, i, gv)
    M0 = np.zeros((N, NCH))
    M1 = np.zeros((N, NCH, 3))
    M2 = np.zeros((N, NCH, 3, 3))
    np.add.at(M0, i, w)
    np.add.at(M1, i, w[:, :, None] * u[:, None, :])
    np.add.at(M2, i, w[:, :, None, None] * u[:, None, :, None] * u[:, None, None, :])
    a0 = M0 ** 2
    a1 = (M1 ** 2).sum(-1)
    a2 = (M2 ** 2).sum((-1, -2))
    A = np.stack([a0 + a1, a0 - a1, a0 + 2 * a1 + a2, a0 - 2 * a1 + a2], -1).reshape(N, 4 * NCH)
    return np.concatenate([G, A], 1), (M0, M1, M2)


def angular_bruteforce(x, L, rc):
    """The 32 angular features as explicit sums over neighbour pairs (l, m)
    of w_c(r_il) w_c(r_im) (1 + lambda cos theta_lim)^zeta (l = m included)."""
    g = geometry(rc)
    N = len(x)
    i, l, rv, r = pairs(x, L, rc)
    w, _ = angular_basis(r, g)
    out = np.zeros((N, NCH, 4))
    for b in range(N):
        sel = i == b
        u = rv[sel] / r[sel][:, None]
        wb = w[sel]
        cos = u @ u.T
        for c in range(NCH):
            ww = np.outer(wb[:, c], wb[:, c])
            out[b, c, 0] = (ww * (1 + cos)).sum()
            out[b, c, 1] = (ww * (1 - cos)).sum()
            out[b, c, 2] = (ww * (1 + cos) ** 2).sum()
            out[b, c, 3] = (ww * (1 - cos) ** 2).sum()
    return out.reshape(N, 4 * NCH)


def network(Dn, par):
    W1, b1, W2, b2, w3, b3 = split_params(par)
    z1 = Dn @ W1 + b1
    h1 = np.tanh(z1)
    z2 = h1 @ W2 + b2
    h2 = np.tanh(z2)
    y = h2 @ w3 + b3
    return y, (h1, h2)


def energies(x, L, rc, par, norm):
    """Per-particle predicted energies e_i (Hartree) of one slice."""
    mean, istd, ymu, ysd = norm
    D, _ = descriptors(x, L, rc)
    y, _ = network((D - mean) * istd, par)
    return y * ysd + ymu


def forces(x, L, rc, par, norm):
    """-d(sum_i e_i)/dx of one slice, analytically through the network and
    the descriptors."""
    mean, istd, ymu, ysd = norm
    g = geometry(rc)
    N = len(x)
    D, (M0, M1, M2) = descriptors(x, L, rc)
    W1, b1, W2, b2, w3, b3 = split_params(par)
    y, (h1, h2) = network((D - mean) * istd, par)
    d2 = w3 * (1 - h2 ** 2)
    d1 = (d2 @ W2.T) * (1 - h1 ** 2)
    dD = (d1 @ W1.T) * istd * ysd  # de_i / dD_i
    gam = dD[:, :NR]
    dA = dD[:, NR:].reshape(N, NCH, 4)
    z0 = 2 * M0 * dA.sum(-1)
    z1 = 2 * M1 * (dA[..., 0] - dA[..., 1] + 2 * dA[..., 2] - 2 * dA[..., 3])[..., None]
    z2 = 2 * M2 * (dA[..., 2] + dA[..., 3])[..., None, None]
    i, l, rv, r = pairs(x, L, rc)
    u = rv / r[:, None]
    _, dg = radial_basis(r, g)
    w, dw = angular_basis(r, g)
    rad = (gam[i] * dg).sum(1)
    t1 = np.einsum("pcd,pd->pc", z1[i], u)
    q = np.einsum("pcab,pb->pca", z2[i], u)
    t2 = np.einsum("pca,pa->pc", q, u)
    rad += (dw * (z0[i] + t1 + t2)).sum(1)
    perp = ((w / r[:, None])[..., None] * ((z1[i] - t1[..., None] * u[:, None, :])
                                          + 2 * (q - t2[..., None] * u[:, None, :]))).sum(1)
    grad = rad[:, None] * u + perp  # d e_i / d r_il, r_il = x_l - x_i
    F = np.zeros((N, 3))
    np.add.at(F, i, grad)
    np.add.at(F, l, -grad)
    return F


def loss_and_gradient(Dn, t, par):
    """Training loss (1/n) sum (y - t)^2 and its gradient w.r.t. the flat
    parameter vector."""
    W1, b1, W2, b2, w3, b3 = split_params(par)
    y, (h1, h2) = network(Dn, par)
    n = len(t)
    e = y - t
    loss = (e ** 2).mean()
    dy = 2 * e / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = dy[:, None] * w3 * (1 - h2 ** 2)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (1 - h1 ** 2)
    gW1 = Dn.T @ d1
    gb1 = d1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
