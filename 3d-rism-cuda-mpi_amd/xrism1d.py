"""1D XRISM solvent susceptibility generator: writes the .xvk files the 3D
solve reads (the format of solvent_read.cc:31-60).

The reference only consumes .xvk files; this module produces them.  For the
solvent sites (LJ sigma/eps, charge, in-molecule coordinates, molar density
per species) it solves the site-site (extended RISM) Ornstein-Zernike
equation of the pure solvent,

    H(k) = W(k) C(k) W(k) + W(k) C(k) P H(k),   P = diag(rho_a),

with W the intramolecular matrix (delta_ab on the diagonal, j0(k l_ab) for
two sites of one molecule, 0 across molecules), and the KH (or HNC) closure
in r, and returns chi_ab(k) = w_ab(k) + rho_b h_ab(k) on the reference's k
grid k_i = (i + 1) pi / (dr ntab), summed over symmetry-equivalent sites b
(the merged sites the .xvk stores).  The Coulomb tail is split off as
beta q_a q_b erf(r/eta)/r and transformed analytically (its transform is
4 pi beta q_a q_b exp(-k^2 eta^2 / 4) / k^2), so the numerical transforms
(sine transforms on r_j = j dr) only see short-range functions.  The
iteration is Anderson mixing with restarts; ionic species are charged up in
four stages.  The isothermal compressibility written to the header is the
Kirkwood-Buff value over all species (one site per species), extrapolated to
k = 0.

Regenerating the bundled example/h2o_vv.xvk, naclaq_vv.xvk, kclaq_vv.xvk and
example_milc/ctip3p.xvk from their own headers reproduces their chi(k) to
2e-7 (1.3e-5 for ctip3p) and their compressibility to 2e-5 relative, which
shows those files were made by XRISM-KH on the same grids.  With a target
dielectric constant the solve is the dielectrically consistent DRISM
(`dielectric=`); the dielectric constant of any solution follows from the
k -> 0 charge sum rule (`dielectric_constant`): 1 + 3y for XRISM, the target
for DRISM.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.fft as sfft
from scipy.special import erfc

AVOGADRO = 6.02214076e23
BOLTZMANN = 1.380649e-23
HARTREE = 4.359744722207185e-18
BOHR = 0.5291771500001
KCAL = 4184.0


@dataclass
class Site:
    name: str
    species: int     # molecule type (SPC column)
    sym: int         # symmetry class; a negative value marks a copy of class |sym|
    sig: float       # Angstrom
    eps: float       # J/mol
    q: float         # e
    xyz: tuple       # in-molecule coordinates, Angstrom
    density: float   # mol/L of the species


@dataclass
class Result:
    k: np.ndarray            # (ntab,)
    chi: np.ndarray          # (ntab, natv, natv): chi[i][v2][v1] as the .xvk stores it
    xt: float                # isothermal compressibility, 1/Pa
    iterations: int
    rms: float
    sites: list = field(default_factory=list)


def water_tip3p(density=55.33296, h_sigma=0.4, h_eps_kcal=0.046):
    """TIP3P-type water: O sigma 3.1507 A, eps 0.1521 kcal/mol, q -0.834; H
    sigma / eps as given, q +0.417; OH 0.9572 A, HOH 104.52 deg, the dipole
    along z."""
    half = math.radians(104.52) / 2
    y, z = 0.9572 * math.sin(half), 0.9572 * math.cos(half)
    return [Site("O", 1, 1, 3.1507, 0.1521 * KCAL, -0.834, (0.0, 0.0, 0.0), density),
            Site("H", 1, 2, h_sigma, h_eps_kcal * KCAL, 0.417, (0.0, y, z), density),
            Site("H", 1, -2, h_sigma, h_eps_kcal * KCAL, 0.417, (0.0, -y, z), density)]


def ion(name, species, sym, sig, eps_kcal, q, molar):
    return Site(name, species, sym, sig, eps_kcal * KCAL, q, (0.0, 0.0, 0.0), molar)


def read_header(path):
    """Sites, temperature and grid of an existing .xvk (solvent_read.cc:31-48)."""
    with open(path) as f:
        lines = [next(f) for _ in range(3)]
        h1, h2 = lines[1].split(), lines[2].split()
        num, ntab, dr = int(h1[1]), int(h1[3]), float(h1[4])
        temper = float(h2[2])
        next(f)
        sites = []
        for _ in range(num):
            p = next(f).split()
            sites.append(Site(p[1], int(p[2]), int(p[3]), float(p[4]), float(p[5]), float(p[6]),
                              (float(p[7]), float(p[8]), float(p[9])), float(p[10])))
    return sites, temper, dr, ntab


class _Grid:
    def __init__(self, dr, ntab):
        self.n = ntab - 1
        self.dr = dr
        self.dk = math.pi / (ntab * dr)
        self.r = dr * np.arange(1, self.n + 1)
        self.k = self.dk * np.arange(1, self.n + 1)

    # f(k) = 4 pi / k int r f(r) sin(k r) dr; f(r) = 1/(2 pi^2 r) int k F sin(k r) dk
    def fwd(self, f):
        return 2.0 * math.pi * self.dr / self.k * sfft.dst(self.r * f, type=1, axis=-1)

    def inv(self, F):
        return self.dk / (4.0 * math.pi ** 2 * self.r) * sfft.dst(self.k * F, type=1, axis=-1)


def _closure(name, e):
    if name == "HNC":
        return np.expm1(e)
    if name == "KH":
        return np.where(e > 0, e, np.expm1(np.minimum(e, 0.0)))
    raise ValueError(f"unknown closure {name}")


def _drism_zeta(sites, k, ub, dielectric, a_smooth):
    """The dielectric bridge of DRISM (Perkyns & Pettitt, Chem. Phys. Lett. 190,
    626 (1992)) for the one polar species: zeta_ij(k) = f_i(k) f_j(k) h_c(k),
    f_i = j0(k x_i) j0(k y_i) j1(k z_i) in the molecule's frame (centre of its
    sites at the origin, dipole along z), h_c(k) = h_c0 exp(-k^2 a^2 / 4),
    y = 4 pi beta rho mu^2 / 9.  Its k -> 0 part adds rho y h_c0 to the
    dielectric constant XRISM gives (1 + 3 y), so h_c0 = ((eps - 1) / y - 3) /
    rho makes the charge sum rule of the solution give eps itself (the form
    (eps / y - 3) / rho realises eps + 1 under the same sum rule).  Returns
    (zeta, rho) with zeta an (ns, ns, n) array."""
    from scipy.special import spherical_jn

    ns = len(sites)
    zeta = np.zeros((ns, ns, len(k)))
    polar = {}
    for i, s in enumerate(sites):
        polar.setdefault(s.species, []).append(i)
    mus = {}
    for spc, idx in polar.items():
        if len(idx) < 2:
            continue
        xyz = np.array([sites[i].xyz for i in idx], dtype=float)
        qv = np.array([sites[i].q for i in idx])
        mu = (qv[:, None] * xyz).sum(0)
        if np.linalg.norm(mu) > 1e-12:
            mus[spc] = (idx, xyz, mu)
    if len(mus) != 1:
        raise ValueError("DRISM here supports exactly one polar species")
    (spc, (idx, xyz, mu)), = mus.items()
    rho = sites[idx[0]].density * AVOGADRO * 1e-27
    # rotate the dipole onto +z
    zhat = mu / np.linalg.norm(mu)
    v = np.cross(zhat, [0.0, 0.0, 1.0])
    c = float(np.dot(zhat, [0.0, 0.0, 1.0]))
    if np.linalg.norm(v) < 1e-12:
        rot = np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    else:
        vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
        rot = np.eye(3) + vx + vx @ vx * (1.0 / (1.0 + c))
    loc = (xyz - xyz.mean(0)) @ rot.T
    mu2 = float(np.dot(mu, mu))
    y = 4.0 * math.pi * ub * rho * mu2 / 9.0
    hc = ((dielectric - 1.0) / y - 3.0) / rho * np.exp(-(k * a_smooth) ** 2 / 4.0)
    f = [spherical_jn(0, k * p[0]) * spherical_jn(0, k * p[1]) * spherical_jn(1, k * p[2])
         for p in loc]
    for a_, i in enumerate(idx):
        for b_, j in enumerate(idx):
            zeta[i, j] = f[a_] * f[b_] * hc
    return zeta, rho


def dielectric_constant(res, sites, temper):
    """eps from the k -> 0 limit of the charge structure factor, 1 - 1/eps =
    (4 pi beta / k^2) sum_ab q_a q_b rho_a chi_ab(k), over the two smallest k
    extrapolated quadratically (the merged chi summed over the sites a)."""
    ub = HARTREE * BOHR / (BOLTZMANN * temper)
    classes, mult, qc, rc = [], {}, {}, {}
    for s in sites:
        c = abs(s.sym)
        if c not in classes:
            classes.append(c)
            qc[c], rc[c] = s.q, s.density * AVOGADRO * 1e-27
        mult[c] = mult.get(c, 0) + 1
    vals = []
    for kk in (0, 1):
        sqq = sum(mult[a] * rc[a] * qc[a] * qc[b] * res.chi[kk, i, j]
                  for i, a in enumerate(classes) for j, b in enumerate(classes))
        vals.append(4.0 * math.pi * ub * sqq / res.k[kk] ** 2)
    x = vals[0] + (vals[0] - vals[1]) / 3.0
    return 1.0 / (1.0 - x)


def solve(sites, temper, dr=0.05, ntab=8192, closure="KH", tol=1e-11, eta=1.0, mix=0.3,
          history=10, maxiter=3000, stages=(0.25, 0.5, 0.75, 1.0), dielectric=None,
          drism_a=1.0):
    """Solves XRISM for the given sites (DRISM with `dielectric` = the target
    dielectric constant); returns a Result."""
    g = _Grid(dr, ntab)
    ns, n = len(sites), g.n
    kT = BOLTZMANN * temper
    rt = AVOGADRO * kT
    ub = HARTREE * BOHR / kT  # beta e^2 / (4 pi eps0), Angstrom (cal_Coulomb.cu:33)
    sig = np.array([s.sig for s in sites])
    eps = np.array([s.eps for s in sites])
    q = np.array([s.q for s in sites])
    rho = np.array([s.density for s in sites]) * AVOGADRO * 1e-27  # 1/A^3
    xyz = np.array([s.xyz for s in sites], dtype=float)
    W = np.zeros((ns, ns, n))
    for a in range(ns):
        for b in range(ns):
            if sites[a].species != sites[b].species:
                continue
            if a == b:
                W[a, b] = 1.0
            else:
                l = float(np.linalg.norm(xyz[a] - xyz[b]))
                W[a, b] = np.sin(g.k * l) / (g.k * l)
    lj = np.zeros((ns, ns, n))
    qq = q[:, None] * q[None, :]
    for a in range(ns):
        for b in range(ns):
            s = 0.5 * (sig[a] + sig[b])
            e = math.sqrt(eps[a] * eps[b]) / rt
            x6 = (s / g.r) ** 6
            lj[a, b] = 4.0 * e * (x6 * x6 - x6)
    short = ub * qq[:, :, None] * erfc(g.r / eta) / g.r
    longk = ub * qq[:, :, None] * 4.0 * math.pi / g.k ** 2 * np.exp(-(g.k * eta) ** 2 / 4.0)
    zeta = None
    if dielectric is not None:
        zeta, rho_p = _drism_zeta(sites, g.k, ub, dielectric, drism_a)
        W2 = W + rho_p * zeta  # omega + rho zeta
    else:
        W2 = W
    Wk = np.moveaxis(W2, -1, 0)
    eye = np.eye(ns)[None]

    def step(ts, scale):
        h = _closure(closure, -(lj + scale * short) + ts)
        cs = h - ts
        Cs = g.fwd(cs)
        Ck = np.moveaxis(Cs - scale * longk, -1, 0)
        WC = Wk @ Ck
        # (h - zeta) = W2 C W2 + W2 C P (h - zeta)  (DRISM; zeta = 0: XRISM)
        H = np.linalg.solve(eye - WC * rho[None, None, :], WC @ Wk)
        Hk = np.moveaxis(H, 0, -1)
        if zeta is not None:
            Hk = Hk + zeta
        return g.inv(Hk - Cs) - ts, Hk

    ts = np.zeros((ns, ns, n))
    charged = bool(np.any(q != 0))
    total = 0
    rms = np.inf
    Hk = None
    for scale in (stages if charged else (1.0,)):
        R_hist, P_hist = [], []
        best, beta = (np.inf, ts.copy()), mix
        for it in range(maxiter):
            R, Hk = step(ts, scale)
            total += 1
            rms = float(np.sqrt(np.mean(R * R)))
            if not np.isfinite(rms) or rms > 100.0 * best[0]:
                ts, R_hist, P_hist, beta = best[1].copy(), [], [], beta * 0.5
                continue
            if rms < best[0]:
                best = (rms, ts.copy())
            if rms < tol:
                break
            Q = ts + beta * R
            if R_hist and it > 10:
                M = np.stack([(R - Rp).ravel() for Rp in R_hist], 1)
                th = np.linalg.lstsq(M, R.ravel(), rcond=None)[0]
                tn = Q + sum(th[i] * (P_hist[i] - Q) for i in range(len(R_hist)))
            else:
                tn = Q
            R_hist.append(R)
            P_hist.append(Q)
            if len(R_hist) > history:
                R_hist.pop(0)
                P_hist.pop(0)
            ts = tn
        if rms >= tol:
            raise RuntimeError(f"XRISM did not converge (rms {rms:.3e} at charge scale {scale})")
    X = W + Hk * rho[None, :, None]
    # merge symmetry copies: chi[v][v'] = sum over sites b of class v' of X[rep(v), b]
    classes = []
    for s in sites:
        if abs(s.sym) not in classes:
            classes.append(abs(s.sym))
    rep = {c: [abs(s.sym) for s in sites].index(c) for c in classes}
    nv = len(classes)
    chi = np.zeros((ntab, nv, nv))
    for i, ci in enumerate(classes):
        for j, cj in enumerate(classes):
            acc = sum(X[rep[ci], b] for b in range(ns) if abs(sites[b].sym) == cj)
            chi[:n, i, j] = acc
            # k = pi / dr (the last tabulated k): the sine transforms vanish there
            chi[n, i, j] = sum(_w_at(sites, rep[ci], b, ntab * g.dk)
                               for b in range(ns) if abs(sites[b].sym) == cj)
    xt = _compressibility(sites, rho, Hk, kT)
    k = g.dk * np.arange(1, ntab + 1)
    return Result(k, chi, xt, total, rms, list(sites))


def _w_at(sites, a, b, kk):
    if sites[a].species != sites[b].species:
        return 0.0
    if a == b:
        return 1.0
    l = float(np.linalg.norm(np.asarray(sites[a].xyz) - np.asarray(sites[b].xyz)))
    return math.sin(kk * l) / (kk * l)


def _compressibility(sites, rho, Hk, kT):
    """Kirkwood-Buff: 1 / (kT sum_ij rho_i rho_j (B^-1)_ij), B_ij = rho_i delta_ij +
    rho_i rho_j h_ij(k), one site per species, at the two smallest k and
    extrapolated quadratically to k = 0."""
    species = []
    for i, s in enumerate(sites):
        if s.species not in [sites[j].species for j in species]:
            species.append(i)
    rr = np.array([rho[i] for i in species])
    vals = []
    for kk in (0, 1):
        G = np.array([[Hk[a, b][kk] for b in species] for a in species])
        B = np.diag(rr) + rr[:, None] * rr[None, :] * G
        vals.append(1.0 / float(np.sum(rr[:, None] * rr[None, :] * np.linalg.inv(B))) / (1e30 * kT))
    return vals[0] + (vals[0] - vals[1]) / 3.0


def _fortran_e(x):
    """0.ddddddddE+xx in 16 columns, as the bundled headers print it."""
    if x == 0:
        return "0.00000000E+00".rjust(16)
    e = int(math.floor(math.log10(abs(x)))) + 1
    m = x / 10.0 ** e
    if round(abs(m), 8) >= 1.0:
        m /= 10.0
        e += 1
    return f"{m:.8f}E{e:+03d}".rjust(16)


def write_xvk(path, res: Result, temper, dr):
    """Writes res in the .xvk layout (solvent_read.cc:31-60)."""
    ntab, nv = res.chi.shape[0], res.chi.shape[1]
    nspec = len(set(s.species for s in res.sites))
    with open(path, "w") as f:
        f.write("##  reduced xvv(k) data with solvent parameters\n")
        f.write(f"##{len(res.sites):10d}{nv:8d}{ntab:8d}{dr:16.8f}{math.pi / (dr * ntab):16.8f}\n")
        f.write(f"##{nspec:5d}{temper:17.8f}{_fortran_e(res.xt)}\n")
        f.write("## ATOM SPC SYM sig[Angs]   eps[J/mol]  charge[e]    ---X---     ---Y---     "
                "---Z---    density[M] \n")
        for s in res.sites:
            f.write(f"## {s.name:<4s}{s.species:4d}{s.sym:4d}{s.sig:12.5f}{s.eps:12.5f}{s.q:12.5f}"
                    f"{s.xyz[0]:12.5f}{s.xyz[1]:12.5f}{s.xyz[2]:12.5f}{s.density:12.5f}\n")
        f.write("##\n")
        for i in range(ntab):
            f.write(" ".join(f"{v: .15E}" for v in res.chi[i].ravel()) + "\n")


def regenerate(src, dst=None, closure="KH"):
    """Solves for the sites, temperature and grid of an existing .xvk and
    (optionally) writes the result; returns the Result."""
    sites, temper, dr, ntab = read_header(src)
    res = solve(sites, temper, dr, ntab, closure)
    if dst:
        write_xvk(dst, res, temper, dr)
    return res


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="write a .xvk from an XRISM solve")
    ap.add_argument("out")
    ap.add_argument("--like", help="take sites, temperature and grid from this .xvk")
    ap.add_argument("--salt", choices=["none", "NaCl", "KCl"], default="none")
    ap.add_argument("--molar", type=float, default=0.2)
    ap.add_argument("--temper", type=float, default=298.15)
    ap.add_argument("--dr", type=float, default=0.05)
    ap.add_argument("--ntab", type=int, default=8192)
    ap.add_argument("--closure", default="KH")
    ap.add_argument("--dielectric", type=float, default=None,
                    help="DRISM with this dielectric constant (default: XRISM)")
    a = ap.parse_args()
    if a.like:
        regenerate(a.like, a.out, a.closure)
    else:
        sites = water_tip3p(density=0.03334 / AVOGADRO * 1e27)
        cl = ion("Cl", 3, 4, 4.478, 0.0356, -1.0, a.molar)
        if a.salt == "NaCl":
            sites += [ion("Na", 2, 3, 2.439, 0.0874, 1.0, a.molar), cl]
        elif a.salt == "KCl":
            sites += [ion("K", 2, 3, 3.038, 0.1937, 1.0, a.molar), cl]
        res = solve(sites, a.temper, a.dr, a.ntab, a.closure, dielectric=a.dielectric)
        write_xvk(a.out, res, a.temper, a.dr)
