"""Independent numpy model of the separation likelihood, written from the
mathematics rather than from the reference code, to pin the oracle.

It deliberately uses different formulations from the reference (and the
oracle) wherever the mathematics allows, so that a transcription mistake in
the oracle (a dropped term, a sign, a swapped operand) shows up as a
disagreement instead of being copied:

* Gauss-Legendre nodes and weights from numpy.polynomial.legendre.leggauss;
* convolution distances r_i = r' * 10**(dx_i / 5) (the closed form of
  pow(10, (g' + dx_i - 4.2)/5 + 1)/1000 with g' = 5 log10(1000 r') - 5 + 4.2);
* nu-slab weights sin(nu + dnu) - sin(nu) (= cos(90 - nu - dnu) - cos(90 - nu));
* survey geometry as a product of rotation matrices;
* stream centres with x = r cos b cos l - 8.5 (the reference's cancelling
  form is mathematically equal);
* squared distance to the stream axis as |(p - c) x a|^2;
* sums with numpy's pairwise summation.

Test infrastructure only.
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial.legendre import leggauss

SIGMA_MAG = 0.6
ABS_MAG = 4.2
SUN_R0 = 8.5
SIGMOID = (0.9402, 1.6171, 23.5877)
GAL_MATRIX = np.array([[-0.054875539726, -0.873437108010, -0.483834985808],
                       [0.494109453312, -0.444829589425, 0.746982251810],
                       [-0.867666135858, -0.198076386122, 0.455983795705]])
D2R = np.pi / 180.0


def rot_x(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def rot_z(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def wedge_inclination(wedge: int) -> float:
    eta = 2.5 * wedge - 57.5 - (180.0 if wedge > 46 else 0.0)
    return eta + 32.5


def sdss_gc_to_lb(wedge: int, mu, nu):
    """(mu, nu) on the stripe's great circle (node RA 95 deg) -> (l, b) deg."""
    mu = np.asarray(mu, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    node = 95.0 * D2R
    v = np.stack([np.cos(nu * D2R) * np.cos(mu * D2R - node),
                  np.cos(nu * D2R) * np.sin(mu * D2R - node),
                  np.sin(nu * D2R)])
    eq = rot_z(node) @ rot_x(wedge_inclination(wedge) * D2R) @ v.reshape(3, -1)
    g = GAL_MATRIX @ eq
    # spherical latitude atan2(z, rho): the tabulated matrix is orthonormal
    # only to ~1e-12, so the vector is not exactly unit length
    l = np.mod(np.arctan2(g[1], g[0]) / D2R, 360.0)
    b = np.arctan2(g[2], np.hypot(g[0], g[1])) / D2R
    return l.reshape(mu.shape), b.reshape(mu.shape)


def sgr_rotation():
    """Majewski et al. (2003) galactic -> Sagittarius rotation, ZXZ Euler
    angles (phi, theta, psi) = (183.75, 76.54, 194.111534) deg."""
    phi, theta, psi = 183.75 * D2R, (90 - 13.46) * D2R, (180 + 14.111534) * D2R
    return rot_z(-psi) @ rot_x(-theta) @ rot_z(-phi)


def sgr_gc_to_lb(wedge: int, mu, nu):
    """Sagittarius stripes: lambda = atan2(-sin nu, cos mu cos nu) + 2.5 wedge,
    beta = asin(sin mu cos nu); then the inverse Majewski rotation with the
    reference's axis convention (sgr unit vector (cos b cos l, cos b sin l, -sin b))."""
    mu = np.asarray(mu, dtype=np.float64) * D2R
    nu = np.asarray(nu, dtype=np.float64) * D2R
    lam = np.arctan2(-np.sin(nu), np.cos(mu) * np.cos(nu)) + 2.5 * wedge * D2R
    beta = np.arcsin(np.sin(mu) * np.cos(nu))
    v = np.stack([np.cos(beta) * np.cos(lam), np.cos(beta) * np.sin(lam), -np.sin(beta)])
    g = sgr_rotation().T @ v.reshape(3, -1)
    l = np.mod(np.arctan2(g[1], g[0]) / D2R, 360.0)
    b = np.arcsin(np.clip(g[2], -1, 1)) / D2R
    return l.reshape(mu.shape), b.reshape(mu.shape)


def gc_to_lb(wedge, sgr, mu, nu):
    return sgr_gc_to_lb(wedge, mu, nu) if sgr else sdss_gc_to_lb(wedge, mu, nu)


def unit_los(l, b):
    l = np.asarray(l) * D2R
    b = np.asarray(b) * D2R
    return np.stack([np.cos(b) * np.cos(l), np.cos(b) * np.sin(l), np.sin(b)], axis=-1)


def distance_tables(r_prime, n):
    """Per distance r' (kpc, any shape): rho = reff * 3 sigma / r'^3, the
    convolution distances r_i, weights qw_i and magnitudes g_i."""
    x, w = leggauss(n)
    dx = 3 * SIGMA_MAG * x
    rp = np.asarray(r_prime, dtype=np.float64)[..., None]
    gp = 5 * np.log10(1000 * rp) - 5 + ABS_MAG
    reff = SIGMOID[0] / (np.exp(SIGMOID[1] * (gp - SIGMOID[2])) + 1)
    rho = (reff * 3 * SIGMA_MAG / rp ** 3)[..., 0]
    ri = rp * 10 ** (dx / 5)
    coeff = 1 / (SIGMA_MAG * np.sqrt(2 * np.pi))
    qw = w * ri ** 3 * coeff * np.exp(-dx ** 2 / (2 * SIGMA_MAG ** 2))
    return rho, ri, qw, gp + dx


def region_tables(reg, n):
    k = np.arange(reg.r_steps)
    dg = (reg.r_max - reg.r_min) / reg.r_steps
    g = reg.r_min + k * dg
    r0 = 10 ** ((g - 14.2) / 5)
    r1 = 10 ** ((g + dg - 14.2) / 5)
    dmu = (reg.mu_max - reg.mu_min) / reg.mu_steps
    irv = (r1 ** 3 - r0 ** 3) / 3 * dmu * D2R
    rho, ri, qw, gi = distance_tables((r0 + r1) / 2, n)
    dnu = (reg.nu_max - reg.nu_min) / reg.nu_steps
    nu = reg.nu_min + np.arange(reg.nu_steps) * dnu
    ids = np.sin((nu + dnu) * D2R) - np.sin(nu * D2R)
    return dict(irv=irv, rho=rho, ri=ri, qw=qw, gi=gi, ids=ids, nu_mid=nu + dnu / 2,
                mu_mid=reg.mu_min + (np.arange(reg.mu_steps) + 0.5) * dmu)


class Model:
    """Model parameters of an AstronomyParameters object."""

    def __init__(self, p):
        bp = p.background_parameters
        self.alpha, self.q, self.r0, self.delta = bp[0], bp[1], bp[2], bp[3]
        self.aux = p.aux_bg_profile == 1
        self.abc = tuple(bp[4:7]) if self.aux else (0.0, 0.0, 0.0)
        self.centres, self.axes, self.sigmas = [], [], []
        for s in p.streams:
            mu, r, th, ph, sig = s.parameters
            l, b = gc_to_lb(p.wedge, p.sgr_coordinates, np.array([mu]), np.array([0.0]))
            self.centres.append(r * unit_los(l[0], b[0]) - np.array([SUN_R0, 0.0, 0.0]))
            self.axes.append(np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]))
            self.sigmas.append(sig)
        self.wb = np.exp(p.background_weight)
        self.ws = np.array([np.exp(s.weight) for s in p.streams])

    def densities(self, u, ri, qw, gi):
        """u: (..., 3) unit lines of sight; ri, qw, gi: (..., n).  Returns the
        convolved background and per-stream sums sum_i qw_i f(p_i)."""
        p = ri[..., None] * u[..., None, :] - np.array([SUN_R0, 0.0, 0.0])
        if self.q == 0:
            bg = None
        else:
            rg = np.sqrt(p[..., 0] ** 2 + p[..., 1] ** 2 + (p[..., 2] / self.q) ** 2)
            if self.alpha == 1 and self.delta == 1:
                f = 1 / (rg * (rg + self.r0) ** 3)
                if self.aux:
                    a, b, c = self.abc
                    f = f + a * gi ** 2 + b * gi + c
            else:
                f = 1 / (rg ** self.alpha * (rg + self.r0) ** (3 - self.alpha + self.delta))
            bg = np.sum(qw * f, axis=-1)
        st = []
        for c, a, sig in zip(self.centres, self.axes, self.sigmas):
            if abs(sig) < 1e-4:
                st.append(np.zeros(ri.shape[:-1]))
                continue
            d2 = np.sum(np.cross(p - c, a) ** 2, axis=-1)
            st.append(np.sum(qw * np.exp(-d2 / (2 * sig ** 2)), axis=-1))
        return bg, st


def region_integral(p, model: Model, reg):
    t = region_tables(reg, p.convolve)
    n_s = len(p.streams)
    bg_tot, st_tot = [], [[] for _ in range(n_s)]
    mu_g, nu_g = np.meshgrid(t["mu_mid"], t["nu_mid"], indexing="ij")
    l, b = gc_to_lb(p.wedge, p.sgr_coordinates, mu_g, nu_g)
    u = unit_los(l, b)                                   # (M, NV, 3)
    for i in range(reg.mu_steps):
        uu = u[i][:, None, :]                            # (NV, 1, 3)
        bg, st = model.densities(uu, t["ri"][None], t["qw"][None], t["gi"][None])  # (NV, R)
        vol = t["ids"][:, None] * t["irv"][None, :]      # (NV, R)
        if bg is None:
            bg_tot.append(np.sum(-vol))
        else:
            bg_tot.append(np.sum(bg * t["rho"][None, :] * vol))
        for s in range(n_s):
            st_tot[s].append(np.sum(st[s] * t["rho"][None, :] * vol))
    return float(np.sum(bg_tot)), np.array([np.sum(x) for x in st_tot])


def integrals(p, model: Model = None):
    model = model or Model(p)
    regs = [region_integral(p, model, g) for g in p.regions]
    bg = regs[0][0] - sum(r[0] for r in regs[1:])
    st = regs[0][1] - sum((r[1] for r in regs[1:]), np.zeros(len(p.streams)))
    return bg, st, regs


def star_log10p(p, stars, I_bg, I_st, model: Model = None):
    """Per-star log10 P (-238 where P == 0)."""
    model = model or Model(p)
    stars = np.asarray(stars, dtype=np.float64)
    rho, ri, qw, gi = distance_tables(stars[:, 2], p.convolve)
    u = unit_los(stars[:, 0], stars[:, 1])
    bg, st = model.densities(u, ri, qw, gi)
    bgp = -np.ones(len(stars)) if bg is None else bg * rho
    num = model.wb * bgp / I_bg
    for s in range(len(p.streams)):
        num = num + model.ws[s] * st[s] * rho / I_st[s]
    P = num / (0.001 * (model.wb + np.sum(model.ws)))
    with np.errstate(divide="ignore"):
        out = np.where(P != 0, np.log10(np.where(P != 0, P, 1.0)), -238.0)
    return out


def evaluate(p, stars):
    bg, st, regs = integrals(p)
    lp = star_log10p(p, stars, bg, st)
    return dict(I_bg=bg, I_st=st, regions=regs, log10p=lp, likelihood=float(np.sum(lp)) / len(stars))


# ---------------------------------------------------------------------------
# Stream separation, from the mathematics
# ---------------------------------------------------------------------------

def sgr_lb(lam_deg, beta_deg):
    """Galactic (l, b) of Sagittarius (lambda, beta) through the rotation
    matrix, with the reference's axis convention (z = -sin beta)."""
    lam, beta = lam_deg * D2R, beta_deg * D2R
    v = np.array([np.cos(beta) * np.cos(lam), np.cos(beta) * np.sin(lam), -np.sin(beta)])
    g = sgr_rotation().T @ v
    return np.arctan2(g[1], g[0]) / D2R, np.arcsin(np.clip(g[2], -1, 1)) / D2R


def stripe_normal(wedge: int, sgr: int) -> np.ndarray:
    """SDSS: the pole (nu = 90) of the stripe's great circle.  Sagittarius:
    the normal of the plane through the directions of (lambda, 0) and
    (lambda, 90), lambda = 2.5 wedge, oriented as their cross product."""
    if sgr:
        a = unit_los(*sgr_lb(2.5 * wedge, 0.0))
        c = unit_los(*sgr_lb(2.5 * wedge, 90.0))
        n = np.cross(a, c)
        return n / np.linalg.norm(n)
    l, b = sdss_gc_to_lb(wedge, np.array([0.0]), np.array([90.0]))
    return unit_los(l[0], b[0])


def rotation_to_z(n) -> np.ndarray:
    """Rodrigues' formula: the rotation about n x z by the angle between n and
    z, R = I + sin(t) K + (1 - cos(t)) K^2 with K the cross-product matrix of
    the unit axis."""
    n = np.asarray(n, dtype=np.float64) / np.linalg.norm(n)
    z = np.array([0.0, 0.0, 1.0])
    axis = np.cross(n, z)
    s = np.linalg.norm(axis)
    k = axis / s
    t = np.arctan2(s, n @ z)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * (K @ K)


def separation_positions(stars, rot) -> np.ndarray:
    """A star at (l, b, r) is drawn at distance g = 5 log10(100 r) + 4.2 - 11
    (the magnitude-like log distance) along its direction, in the rotated frame."""
    stars = np.asarray(stars, dtype=np.float64)
    g = 5.0 * np.log10(100.0 * stars[:, 2]) + 4.2 - 11.0
    return (g[:, None] * unit_los(stars[:, 0], stars[:, 1])) @ rot.T


def membership(p, stars, I_bg, I_st, model: Model = None) -> np.ndarray:
    """Posterior stream-membership probabilities of each star, [N][S]:
    w_s S_s / I_s over (sum_t w_t S_t / I_t + B / I_b), with w_s = exp(weight_s)
    and the background's weight fixed at 1 (the common normalisation of the
    mixture weights cancels)."""
    model = model or Model(p)
    stars = np.asarray(stars, dtype=np.float64)
    rho, ri, qw, gi = distance_tables(stars[:, 2], p.convolve)
    u = unit_los(stars[:, 0], stars[:, 1])
    bg, st = model.densities(u, ri, qw, gi)
    terms = np.stack([model.ws[s] * st[s] * rho / I_st[s] for s in range(len(p.streams))], axis=1)
    bgp = -np.ones(len(stars)) if bg is None else bg * rho
    return terms / (terms.sum(axis=1) + bgp / I_bg)[:, None]
