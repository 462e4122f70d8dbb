"""FP64 CPU oracle for the 3D-RISM solve -- TEST INFRASTRUCTURE ONLY.

This module drives ``liboracle.so`` (``rism_oracle.c``), a plain-C FP64
restatement of the reference algorithm, and restates the reference's input
parsing and ``.xmu`` formatting in Python.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU baseline; the product path
never routes through it.

Reference files restated here:
  * ``.xvk`` reader          -- solvent_read.cc:10-94
  * ``.inp`` reader          -- read_input.cu:21-97, solute.cu:18-54
  * cell quantities          -- cell.cc:8-28
  * k-shells / chi mapping   -- initialize_g.cu:7-50, solvent_spline.cu:7-96
  * potentials               -- cal_potential.cu:3-43, cal_LJ.cu, cal_Coulomb.cu
  * initial t                -- initialize_tuv.cu:19-23
  * loop                     -- iterate.cu:48-82
  * thermodynamics, .xmu     -- cal_exchem.cu, cal_pmv.cu, cal_pressure.cu,
                                output_xmu.cu:7-82

``AndersonM`` is the one piece with no counterpart in the reference: Anderson
mixing over m history pairs, the generalisation of AN2 (anderson.cu) that the
product offers as ``rism_opts.anderson_m``.  It is written from its definition
(a weighted least-squares problem solved with ``numpy.linalg.lstsq`` on the
explicit difference vectors) and pinned by the finite termination of Anderson
mixing on a linear map (tests/test_oracle_iter.py).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

AVOGADRO = 6.02214076e23  # physical.h:5
BOLTZMANN = 1.380649e-23  # physical.h:11
HARTREE = 4.359744722207185e-18  # physical.h:14
BOHR = 0.5291771500001  # physical.h:8
VERSION = "20220301"  # version.h:5

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    """Load liboracle.so (building it with make when it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_dbl, c_sz = ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    L.or_fft3d.argtypes = [_dp, c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_int]
    L.or_fft3d.restype = None
    L.or_kshells.argtypes = [c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, _dp, _ip]
    L.or_kshells.restype = c_int
    L.or_spline_coef.argtypes = [_dp, _dp, c_int, _dp]
    L.or_spline_coef.restype = None
    L.or_splint.argtypes = [_dp, _dp, _dp, c_int, c_dbl]
    L.or_splint.restype = c_dbl
    L.or_chi_shells.argtypes = [_dp, c_int, c_int, c_dbl, _dp, c_int, _dp]
    L.or_chi_shells.restype = c_int
    L.or_lj.argtypes = [c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_int, _dp, _dp, _dp,
                        c_int, _dp, _dp, c_dbl, c_dbl, _dp]
    L.or_lj.restype = None
    L.or_ubeta.argtypes = [c_dbl, c_dbl]
    L.or_ubeta.restype = c_dbl
    L.or_coulomb.argtypes = [c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_int, _dp, _dp,
                             c_dbl, _dp, _dp]
    L.or_coulomb.restype = None
    L.or_fk.argtypes = [c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_int, _dp, _dp, c_dbl,
                        _dp]
    L.or_fk.restype = None
    rows = [c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_int, c_int, c_int, c_int]
    L.or_lj_rows.argtypes = rows + [c_int, _dp, _dp, _dp, c_int, _dp, _dp, c_dbl, c_dbl, _dp]
    L.or_lj_rows.restype = None
    L.or_coulomb_rows.argtypes = rows + [c_int, _dp, _dp, c_dbl, _dp, _dp]
    L.or_coulomb_rows.restype = None
    L.or_fk_rows.argtypes = rows + [c_int, _dp, _dp, c_dbl, _dp]
    L.or_fk_rows.restype = None
    L.or_closure.argtypes = [c_int, c_sz, _dp, _dp, _dp]
    L.or_closure.restype = None
    L.or_calculate.argtypes = [c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_int, c_int, _dp,
                               _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp]
    L.or_calculate.restype = c_dbl
    L.or_an2_create.argtypes = [c_int, c_sz, _dp, c_int, c_dbl, c_dbl]
    L.or_an2_create.restype = ctypes.c_void_p
    L.or_an2_destroy.argtypes = [ctypes.c_void_p]
    L.or_an2_destroy.restype = None
    L.or_an2_step.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.or_an2_step.restype = None
    L.or_an2_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                             _dp, _dp]
    L.or_an2_get.restype = None
    L.or_leg2.argtypes = [_dp, _dp, _dp]
    L.or_leg2.restype = None
    L.or_thermo.argtypes = [c_int, c_sz, c_dbl, c_int, _dp, _dp, _dp, _dp, _dp,
                            ctypes.POINTER(c_dbl)]
    L.or_thermo.restype = None
    L.or_pmv.argtypes = [c_dbl, c_dbl, c_dbl]
    L.or_pmv.restype = c_dbl
    L.or_pressure.argtypes = [c_int, _dp, c_dbl, c_dbl]
    L.or_pressure.restype = c_dbl
    L.or_solve.argtypes = [c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_int, c_int, _dp, _dp,
                           _dp, _dp, _dp, _dp, _ip, c_int, c_dbl, c_dbl, c_dbl, c_int, _dp,
                           _dp, _dp, _dp, _dp, ctypes.POINTER(c_int), c_int]
    L.or_solve.restype = c_int
    L.or_num_threads.argtypes = []
    L.or_num_threads.restype = c_int
    _lib = L
    return L


# --------------------------------------------------------------------------
# input formats
# --------------------------------------------------------------------------

@dataclass
class Solvent:
    """Merged solvent sites and chi_vv(k) table (solvent_read.cc:10-94)."""
    natv: int
    ntab: int
    dr: float
    temper: float
    xt: float
    sig: np.ndarray
    eps: np.ndarray
    q: np.ndarray
    rho: np.ndarray  # 1/A^3
    xvv: np.ndarray  # [ntab][natv(v2)][natv(v1)] = X[v2][v1](k_i)

    @property
    def ttab(self):
        dk = math.pi / (self.dr * self.ntab)
        return dk * (np.arange(self.ntab) + 1.0)


def read_xvk(path) -> Solvent:
    """solvent_read.cc:31-85: header lines 1-2, num site lines, a '##' line,
    then ntab records of natv^2 reals read as a token stream."""
    with open(path) as f:
        lines = f.read().split("\n")
    h1 = lines[1].split()
    num, natv, ntab, dr = int(h1[1]), int(h1[2]), int(h1[3]), float(h1[4])
    h2 = lines[2].split()
    temper, xt = float(h2[2]), float(h2[3])
    sig = np.zeros(natv)
    eps = np.zeros(natv)
    q = np.zeros(natv)
    rho = np.zeros(natv)
    for i in range(num):
        tok = lines[4 + i].split()
        sym = int(tok[3])
        i2 = abs(sym) - 1
        sig[i2] = float(tok[4])
        eps[i2] = float(tok[5])
        q[i2] = float(tok[6])
        rho[i2] += float(tok[10])
    rho *= AVOGADRO * 1.0e-27
    toks = " ".join(lines[num + 5:]).split()
    need = ntab * natv * natv
    vals = np.array([float(t) for t in toks[:need]], dtype=np.float64)
    xvv = vals.reshape(ntab, natv, natv)
    return Solvent(natv, ntab, dr, temper, xt, sig, eps, q, rho, xvv)


@dataclass
class Solute:
    sig: np.ndarray
    eps: np.ndarray
    q: np.ndarray
    xyz: np.ndarray  # [natu][3]
    shift: np.ndarray = field(default_factory=lambda: np.zeros(3))

    @property
    def natu(self):
        return len(self.sig)


@dataclass
class Control:
    outlist: str
    ksave: int
    closure: str  # "KH" or "HNC"
    fsolvent: str
    convergence: float
    maxstep: int
    count: int
    m: float
    mp: float
    box: np.ndarray
    grid: np.ndarray


def _c_round(x):
    """C round(): half away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def centering(xyz):
    """solute.cu:18-48: integer-Angstrom shift that centres the bounding box."""
    mn = xyz.min(axis=0)
    mx = xyz.max(axis=0)
    shift = np.array([_c_round(-(mx[i] - mn[i]) / 2 - mn[i]) for i in range(3)], float)
    return xyz + shift, shift


def read_inp(control_path, structure_path=None, center=True, q0=False):
    """read_input.cu:21-97.  Returns (Control, Solute)."""
    with open(control_path) as f:
        tok = f.read().split()
    pos = 0

    def nxt():
        nonlocal pos
        t = tok[pos]
        pos += 1
        return t

    outlist, ksave, check = nxt(), int(nxt()), nxt()
    if check != VERSION:
        raise ValueError("Error: This input file is for old version.")
    closure = nxt()
    closure_code(closure)
    fsolvent = nxt()
    convergence, maxstep = float(nxt()), int(nxt())
    count, m, mp = int(nxt()), float(nxt()), float(nxt())
    box = np.array([float(nxt()) for _ in range(3)])
    grid = np.array([int(nxt()) for _ in range(3)])
    if structure_path:
        with open(structure_path) as f:
            tok = f.read().split()
        pos = 0
    s = nxt()
    if not s.isdigit():
        raise ValueError("Error: Number of solute atom is not a natural number.")
    natu = int(s)
    arr = np.array([float(nxt()) for _ in range(6 * natu)]).reshape(natu, 6)
    sol = Solute(arr[:, 0].copy(), arr[:, 1].copy(), arr[:, 2].copy(), arr[:, 3:6].copy())
    if center:
        sol.xyz, sol.shift = centering(sol.xyz)
    if q0:
        sol.q = np.zeros_like(sol.q)
    ctl = Control(outlist, ksave, closure, fsolvent, convergence, maxstep, count, m, mp,
                  box, grid)
    return ctl, sol


# --------------------------------------------------------------------------
# the solve
# --------------------------------------------------------------------------

def fft3d(data, box, key):
    """Half-shifted transform of a complex [nz][ny][nx] array (copy)."""
    a = np.ascontiguousarray(data, dtype=np.complex128).copy()
    nz, ny, nx = a.shape
    v = a.view(np.float64).reshape(-1)
    lib().or_fft3d(v, nx, ny, nz, float(box[0]), float(box[1]), float(box[2]), int(key))
    return v.view(np.complex128).reshape(nz, ny, nx)


def kshells(grid, box):
    nx, ny, nz = (int(g) for g in grid)
    N = nx * ny * nz
    ga = np.zeros(N)
    ind = np.zeros(N, dtype=np.int32)
    nga = lib().or_kshells(nx, ny, nz, float(box[0]), float(box[1]), float(box[2]), ga, ind)
    return ga[:nga].copy(), ind


def chi_shells(solv: Solvent, ga):
    chi = np.zeros(len(ga) * solv.natv * solv.natv)
    rc = lib().or_chi_shells(np.ascontiguousarray(solv.xvv.reshape(-1)), solv.natv,
                             solv.ntab, solv.dr, np.ascontiguousarray(ga), len(ga), chi)
    if rc != 0:
        raise ValueError("insufficient maximal T tabulated")
    return chi


def potentials(grid, box, solv: Solvent, sol: Solute, lam=1.0, adswitch=0):
    """cal_potential.cu:3-36.  Returns (u [natv][N], fr [N], fk complex [N])."""
    L = lib()
    nx, ny, nz = (int(g) for g in grid)
    N = nx * ny * nz
    b = [float(x) for x in box]
    xyz = np.ascontiguousarray(sol.xyz.reshape(-1), dtype=np.float64)
    lam1 = lam if adswitch == 1 else 1.0
    u = np.zeros(solv.natv * N)
    L.or_lj(nx, ny, nz, *b, sol.natu, xyz, np.ascontiguousarray(sol.sig),
            np.ascontiguousarray(sol.eps), solv.natv, np.ascontiguousarray(solv.sig),
            np.ascontiguousarray(solv.eps), solv.temper, lam1, u)
    de = np.zeros(N)
    fr = np.zeros(N)
    fk = np.zeros(2 * N)
    if adswitch != 1:
        lam2 = lam if adswitch == 2 else 1.0
        ub = L.or_ubeta(solv.temper, lam2)
        q = np.ascontiguousarray(sol.q, dtype=np.float64)
        L.or_coulomb(nx, ny, nz, *b, sol.natu, xyz, q, ub, de, fr)
        L.or_fk(nx, ny, nz, *b, sol.natu, xyz, q, ub, fk)
    u = u.reshape(solv.natv, N)
    u += solv.q[:, None] * de[None, :]
    return u.reshape(-1).copy(), fr, fk


def potentials_rows(grid, box, solv: Solvent, sol: Solute, z0, nzs, y0, nys):
    """potentials() on the rows z in [z0, z0 + nzs), y in [y0, y0 + nys) of the
    full grid (every x): u [natv][nzs][nys][nx], fr [nzs][nys][nx], and fk on
    the same index block of the full k-grid, complex."""
    L = lib()
    nx, ny, nz = (int(g) for g in grid)
    b = [float(x) for x in box]
    blk = [nx, ny, nz] + b + [int(z0), int(nzs), int(y0), int(nys)]
    n = nx * int(nzs) * int(nys)
    xyz = np.ascontiguousarray(sol.xyz.reshape(-1), dtype=np.float64)
    q = np.ascontiguousarray(sol.q, dtype=np.float64)
    u = np.zeros(solv.natv * n)
    L.or_lj_rows(*blk, sol.natu, xyz, np.ascontiguousarray(sol.sig),
                 np.ascontiguousarray(sol.eps), solv.natv, np.ascontiguousarray(solv.sig),
                 np.ascontiguousarray(solv.eps), solv.temper, 1.0, u)
    de, fr, fk = np.zeros(n), np.zeros(n), np.zeros(2 * n)
    ub = L.or_ubeta(solv.temper, 1.0)
    L.or_coulomb_rows(*blk, sol.natu, xyz, q, ub, de, fr)
    L.or_fk_rows(*blk, sol.natu, xyz, q, ub, fk)
    u = u.reshape(solv.natv, n) + solv.q[:, None] * de[None, :]
    shp = (int(nzs), int(nys), nx)
    return (u.reshape((solv.natv,) + shp), fr.reshape(shp),
            fk.view(np.complex128).reshape(shp))


def cubic_shells(grid, box):
    """Shells of a cubic box in closed form: with odd o = 2 j - N + 1 per axis,
    |k|^2 = (pi/L)^2 (ox^2 + oy^2 + oz^2) = (pi/L)^2 (8 i + 3), so the shell of
    a point is i = (ox^2 + oy^2 + oz^2 - 3) / 8 and ga[i] = (pi/L) sqrt(8 i + 3)
    for every i up to the largest.  Equals kshells() up to the labelling of
    shells (and unrealised i), checked in test_oracle_setup.py; used where
    sorting N values would be slow (512^3)."""
    nx, ny, nz = (int(g) for g in grid)
    assert nx == ny == nz and box[0] == box[1] == box[2]
    o = (2 * np.arange(nx, dtype=np.int64) - nx + 1) ** 2
    s = (o[:, None, None] + o[None, :, None] + o[None, None, :]).reshape(-1)
    indga = ((s - 3) // 8).astype(np.int32)
    imax = int(indga.max())
    ga = (np.pi / float(box[0])) * np.sqrt(8.0 * np.arange(imax + 1) + 3.0)
    return ga, indga


def full_spectrum(half):
    """The full k-grid spectrum [nz][ny][nx] of a real field from its k_z > 0
    half [nz/2][ny][nx]: F(-k) = conj F(k) and k_{N-1-j} = -k_j per axis."""
    m, ny, nx = half.shape
    out = np.empty((2 * m, ny, nx), dtype=np.complex128)
    out[m:] = half
    out[:m] = np.conj(half[::-1, ::-1, ::-1])
    return out


@dataclass
class Result:
    steps: int
    status: int  # 0 converged, 1 diverged, 2 step limit
    rms: np.ndarray
    t: np.ndarray
    g: np.ndarray
    h: np.ndarray
    R: np.ndarray


class OracleRISM:
    """The reference pipeline (initialize -> iterate -> output) on the CPU."""

    def __init__(self, grid, box, solv: Solvent, sol: Solute, closure="KH", count=2,
                 m=0.6, mp=0.2, tol=1e-6, maxstep=10000, lam=1.0, adswitch=0):
        self.grid = np.array(grid, dtype=int)
        self.box = np.array(box, dtype=float)
        self.solv = solv
        self.sol = sol
        self.clos = closure_code(closure)
        self.count, self.m, self.mp = int(count), float(m), float(mp)
        self.tol, self.maxstep = float(tol), int(maxstep)
        self.lam, self.adswitch = lam, adswitch
        nx, ny, nz = self.grid
        self.N = int(nx * ny * nz)
        self.dv = float(np.prod(self.box)) / self.N

    def setup(self):
        self.ga, self.indga = kshells(self.grid, self.box)
        self.chi = chi_shells(self.solv, self.ga)
        self.u, self.fr, self.fk = potentials(self.grid, self.box, self.solv, self.sol,
                                              self.lam, self.adswitch)
        # initialize_tuv.cu:19-23
        self.t0 = (self.solv.q[:, None] * self.fr[None, :]).reshape(-1).copy()
        return self

    def calculate(self, t):
        """One reference step from t: returns (rms, g, R, h)."""
        n = self.solv.natv
        g = np.zeros(n * self.N)
        R = np.zeros(n * self.N)
        h = np.zeros(n * self.N)
        work = np.zeros(4 * n * self.N)
        nx, ny, nz = (int(v) for v in self.grid)
        rms = lib().or_calculate(nx, ny, nz, *[float(b) for b in self.box], n, self.clos,
                                 np.ascontiguousarray(self.solv.q), np.ascontiguousarray(t),
                                 self.u, self.fr, self.fk, self.chi, self.indga, g, R, h,
                                 work)
        return rms, g, R, h

    def solve(self, t=None, verbose=False, maxstep=None):
        n = self.solv.natv
        t = (self.t0 if t is None else t).copy()
        g = np.zeros(n * self.N)
        h = np.zeros(n * self.N)
        R = np.zeros(n * self.N)
        ms = self.maxstep if maxstep is None else int(maxstep)
        hist = np.zeros(max(ms, 1))
        st = ctypes.c_int(0)
        nx, ny, nz = (int(v) for v in self.grid)
        steps = lib().or_solve(nx, ny, nz, *[float(b) for b in self.box], n, self.clos,
                               np.ascontiguousarray(self.solv.q),
                               np.ascontiguousarray(self.solv.rho), self.u, self.fr, self.fk,
                               self.chi, self.indga, self.count, self.m, self.mp, self.tol,
                               ms, t, g, h, R, hist, ctypes.byref(st), int(verbose))
        self.result = Result(steps, st.value, hist[:steps].copy(), t, g, h, R)
        return self.result

    def solve_anderson_m(self, m, beta=None, count=None, t=None, maxstep=None):
        """The same loop (iterate.cu:48-72) with AndersonM in place of AN2."""
        n = self.solv.natv
        t = (self.t0 if t is None else t).copy()
        acc = AndersonM(n, self.N, self.solv.rho, m, self.m if beta is None else beta,
                        self.count if count is None else count)
        ms = self.maxstep if maxstep is None else int(maxstep)
        hist, status = [], 2
        g = h = R = None
        for _ in range(ms):
            rms, g, R, h = self.calculate(t)
            hist.append(rms)
            if not np.isfinite(rms):
                status = 1
                break
            if rms <= self.tol:
                status = 0
                break
            t = acc.step(t, R)
        self.result = Result(len(hist), status, np.array(hist), t, g, h, R)
        self.accel = acc
        return self.result

    def thermo(self, h=None, t=None):
        r = self.result if h is None else None
        h = r.h if h is None else h
        t = r.t if t is None else t
        return thermo(self.solv, self.dv, self.clos, h, t, self.u)


def closure_code(name):
    """KH -> 0, HNC -> 1 (read_input.cu:38-47); PSEn / PSE-n -> 10 + n."""
    if name in ("KH", "HNC"):
        return {"KH": 0, "HNC": 1}[name]
    if name.startswith("PSE"):
        num = name[4:] if name[3:4] == "-" else name[3:]
        if num.isdigit() and 1 <= int(num) <= 20:
            return 10 + int(num)
    raise ValueError("Error: Unexpected closure switch ")


class AndersonM:
    """Anderson mixing over m history pairs (no counterpart in the reference).

    With Q = t + beta R and P_p = t_p + beta R_p of the history pairs, a step
    returns  t' = Q + sum_p th_p (P_p - Q),  where th minimises the weighted
    norm |R - sum_p th_p (R - R_p)|, the weight of site v being 1/rho_v as in
    AN2's cal_theta (anderson.cu:81-102).  The first ``count`` steps are the
    reference's Picard steps t' = t + R (anderson.cu:54-60).  Every step then
    stores (R, Q) and drops the oldest pair beyond m.
    """

    def __init__(self, natv, N, rho, m, beta, count=0):
        self.m, self.beta, self.count = int(m), float(beta), int(count)
        self.w = np.repeat(1.0 / np.asarray(rho, dtype=float), N)
        self.R, self.P = [], []
        self.theta = np.zeros(0)

    def coefficients(self, R):
        if not self.R:
            return np.zeros(0)
        sw = np.sqrt(self.w)
        M = np.stack([(R - Rp) * sw for Rp in self.R], axis=1)
        return np.linalg.lstsq(M, R * sw, rcond=None)[0]

    def step(self, t, R):
        Q = t + self.beta * R
        if self.count > 0:
            self.count -= 1
            tn = t + R
            self.theta = np.zeros(0)
        else:
            self.theta = self.coefficients(R)
            tn = Q.copy()
            for th, Pp in zip(self.theta, self.P):
                tn += th * (Pp - Q)
        self.R.append(R.copy())
        self.P.append(Q)
        if len(self.R) > self.m:
            self.R.pop(0)
            self.P.pop(0)
        return tn


def thermo(solv: Solvent, dv, clos, h, t, u=None):
    """cal_exchem/cal_pmv/cal_pressure -> dict of .xmu quantities (J/mol etc.).

    PSE-n (clos = 10 + n, an extension): the SC integrand is the HNC one minus
    (t*)^(n+1)/(n+1)! where t* = -u + t > 0 (Kast & Kloss 2008); needs u."""
    L = lib()
    n = solv.natv
    N = len(h) // n
    xmu = np.zeros(2 * n)
    xmu2 = np.zeros(n)
    cuv = ctypes.c_double(0.0)
    L.or_thermo(n, N, float(dv), int(clos if clos < 10 else 1),
                np.ascontiguousarray(solv.rho),
                np.ascontiguousarray(h), np.ascontiguousarray(t), xmu, xmu2, ctypes.byref(cuv))
    if clos >= 10:
        order = clos - 10
        e = (np.asarray(t) - np.asarray(u)).reshape(n, N)
        corr = np.where(e > 0, np.maximum(e, 0) ** (order + 1) / math.factorial(order + 1), 0.0)
        xmu[:n] -= corr.sum(axis=1) * solv.rho * dv
    pmv = L.or_pmv(solv.xt, solv.temper, cuv.value)
    pres = L.or_pressure(n, np.ascontiguousarray(solv.rho), solv.xt, solv.temper)
    ibeta = AVOGADRO * BOLTZMANN * solv.temper
    return {
        "SFE_SC": ibeta * xmu[:n].sum(),
        "SFEC_SC": ibeta * xmu[:n],
        "SFE_SC_HNC": ibeta * xmu2.sum() if clos == 0 else None,
        "SFEC_SC_HNC": ibeta * xmu2 if clos == 0 else None,
        "SFE_GF": ibeta * xmu[n:].sum(),
        "SFEC_GF": ibeta * xmu[n:],
        "PMV": pmv * 1000.0,
        "Pressure": ibeta * pres,
        "Correction_Term": -pres * pmv * ibeta,
    }


def atom_sums(kind, grid, box, solv: Solvent, sol: Solute, g, lam=1.0, adswitch=0):
    """Per-atom integrals of g (cal_grad.cu:44-90, cal_euv.cu:30-70,
    cal_ad1.cu:40-80, cal_ad2.cu:40-75), scaled as output_grad.cu,
    output_euv.cu and output_ad.cu write them.  g: [natv][N] on the grid.
    kind 0 grad -> (natu, 3) kcal/mol/A; 1 euv -> (natu, 1 + natv) J/mol
    (total, then per site); 2 ad -> (natu,) kcal/mol (cal_ad1 when
    adswitch == 1, else cal_ad2).  Brute force over all points."""
    nx, ny, nz = (int(v) for v in grid)
    dr = np.asarray(box, dtype=float) / np.array([nx, ny, nz])
    dv = float(np.prod(box)) / (nx * ny * nz)
    cc = HARTREE * BOHR * AVOGADRO
    lam1 = lam if adswitch == 1 else 1.0
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rx = ((x - nx // 2) * dr[0]).reshape(-1)
    ry = ((y - ny // 2) * dr[1]).reshape(-1)
    rz = ((z - nz // 2) * dr[2]).reshape(-1)
    g = np.asarray(g).reshape(solv.natv, -1)
    nu, nv = len(sol.q), solv.natv
    out = []
    for a in range(nu):
        d = (rx - sol.xyz[a, 0], ry - sol.xyz[a, 1], rz - sol.xyz[a, 2])
        r2 = d[0] ** 2 + d[1] ** 2 + d[2] ** 2
        r1 = np.sqrt(r2)
        acc = np.zeros(3 if kind == 0 else (nv if kind == 1 else 1))
        for v in range(nv):
            sg = (sol.sig[a] + solv.sig[v]) * 0.5 * lam1
            eps = math.sqrt(sol.eps[a] * solv.eps[v])
            qq = sol.q[a] * solv.q[v] * cc
            cut = sg * (lam if (kind == 2 and adswitch == 1) else 1.0) * 0.5
            m = r1 >= cut
            with np.errstate(divide="ignore", invalid="ignore"):
                rs6 = (sg * sg / r2[m]) ** 3
                gv = g[v][m]
                if kind == 0:
                    f = (eps * 24.0 * rs6 / r2[m] * (2.0 * rs6 - 1.0) + qq / (r2[m] * r1[m])) * gv
                    acc += solv.rho[v] * np.array([(f * d[i][m]).sum() for i in range(3)])
                elif kind == 1:
                    acc[v] += solv.rho[v] * ((eps * 4.0 * rs6 * (rs6 - 1.0) + qq / r1[m]) * gv).sum()
                elif adswitch == 1:
                    acc[0] += solv.rho[v] * (eps * 24.0 * rs6 * (2.0 * rs6 - 1.0) / lam * gv).sum()
                else:
                    acc[0] += solv.rho[v] * (qq / r1[m] * gv).sum()
        if kind == 1:
            out.append(np.concatenate([[dv * acc.sum()], dv * acc]))
        else:
            out.append(acc * dv / 4184.0)
    out = np.array(out)
    return out[:, 0] if kind == 2 else out


def parse_xmu(path):
    """Reads the total lines of a .xmu file into a dict of floats."""
    out = {}
    with open(path) as f:
        for line in f:
            line = line.strip()
            if "=" not in line or line.startswith("$"):
                continue
            key, rest = line.split("=", 1)
            key = key.strip()
            if "(" in key:
                continue
            out[key] = float(rest.split()[0])
    return out


def run_inp(control_path, structure_path=None, center=True, q0=False, solvent_dir=None,
            verbose=False):
    """Full reference pipeline for a .inp file; returns (oracle, thermo dict)."""
    ctl, sol = read_inp(control_path, structure_path, center, q0)
    sdir = solvent_dir or os.path.dirname(os.path.abspath(control_path))
    solv = read_xvk(os.path.join(sdir, ctl.fsolvent))
    o = OracleRISM(ctl.grid, ctl.box, solv, sol, ctl.closure, ctl.count, ctl.m, ctl.mp,
                   ctl.convergence, ctl.maxstep).setup()
    o.solve(verbose=verbose)
    return o, o.thermo()
