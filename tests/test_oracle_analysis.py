"""The oracle's per-atom analysis integrals (cal_grad, cal_euv, cal_ad1,
cal_ad2), pinned by what the mathematics fixes rather than by retyping the
integrands: grad is the derivative of the g-weighted interaction energy with
respect to the atom's position (central differences), ad1 is the derivative
of the LJ part with respect to the sigma scaling lambda, ad2 is the
Coulomb part of euv, and euv itself reduces to one pair energy (closed
form, CODATA constants) when g is non-zero at a single grid point."""
import math

import numpy as np
import pytest


def _setup(oracle, seed=5):
    rng = np.random.default_rng(seed)
    natv = 2
    solv = oracle.Solvent(natv, 1, 0.1, 300.0, 4.5e-10, np.array([3.15, 0.4]),
                          np.array([636.0, 192.0]), np.array([-0.8, 0.4]),
                          np.array([0.0334, 0.0668]), np.zeros((1, natv, natv)))
    sol = oracle.Solute(np.array([3.4, 2.5]), np.array([360.0, 63.0]), np.array([0.3, -0.2]),
                        np.array([[0.31, -0.27, 0.13], [-1.12, 0.58, 0.44]]))
    grid, box = (14, 12, 16), (7.0, 6.0, 8.0)
    N = int(np.prod(grid))
    # a smooth positive g
    z, y, x = np.meshgrid(*[np.arange(n) for n in grid[::-1]], indexing="ij")
    g = np.stack([1.0 + 0.3 * np.sin(0.4 * x + 0.3 * v) * np.cos(0.5 * y) + 0.1 * z / grid[2]
                  for v in range(natv)]).reshape(natv, N)
    return solv, sol, grid, box, g


def _energy(oracle, solv, sol, grid, box, g, a, dxyz=(0, 0, 0)):
    sol2 = oracle.Solute(sol.sig, sol.eps, sol.q, sol.xyz.copy())
    sol2.xyz[a] += np.asarray(dxyz)
    return oracle.atom_sums(1, grid, box, solv, sol2, g)[a, 0] / 4184.0  # kcal/mol


def test_grad_is_position_derivative(oracle):
    solv, sol, grid, box, g = _setup(oracle)
    gr = oracle.atom_sums(0, grid, box, solv, sol, g)
    h = 1e-5
    for a in range(2):
        for i in range(3):
            e = np.zeros(3)
            e[i] = h
            fd = (_energy(oracle, solv, sol, grid, box, g, a, e) -
                  _energy(oracle, solv, sol, grid, box, g, a, -e)) / (2 * h)
            assert fd == pytest.approx(gr[a, i], rel=1e-5, abs=1e-9)


def test_ad1_is_lambda_derivative(oracle):
    """cal_ad1 is d/dlambda of the LJ energy sum_r 4 eps [(lambda sigma / r)^12 -
    (lambda sigma / r)^6] g over the points it includes (r >= sigma lambda^2 / 2,
    the reference's cut-off for ad1)."""
    solv, sol, grid, box, g = _setup(oracle)
    solv0 = oracle.Solvent(solv.natv, 1, 0.1, 300.0, 4.5e-10, solv.sig, solv.eps,
                           np.zeros(2), solv.rho, solv.xvv)
    lam, h = 0.9, 1e-6
    ad1 = oracle.atom_sums(2, grid, box, solv0, sol, g, lam=lam, adswitch=1)
    nx, ny, nz = grid
    dr = np.asarray(box) / np.array(grid)
    dv = float(np.prod(box)) / (nx * ny * nz)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([((x - nx // 2) * dr[0]).ravel(), ((y - ny // 2) * dr[1]).ravel(),
                    ((z - nz // 2) * dr[2]).ravel()], 1)
    for a in range(2):
        r = np.linalg.norm(pts - sol.xyz[a], axis=1)

        def e_lj(l):
            tot = 0.0
            for v in range(2):
                sg = (sol.sig[a] + solv.sig[v]) * 0.5
                eps = math.sqrt(sol.eps[a] * solv.eps[v])
                m = r >= sg * lam * lam * 0.5
                x6 = (l * sg / r[m]) ** 6
                tot += solv.rho[v] * (4 * eps * (x6 * x6 - x6) * g[v][m]).sum()
            return tot * dv / 4184.0

        fd = (e_lj(lam + h) - e_lj(lam - h)) / (2 * h)
        assert ad1[a] == pytest.approx(fd, rel=1e-6)


def test_ad2_is_coulomb_energy(oracle):
    solv, sol, grid, box, g = _setup(oracle)
    sol_q = oracle.Solute(sol.sig, np.zeros(2), sol.q, sol.xyz)
    solv_q = oracle.Solvent(solv.natv, 1, 0.1, 300.0, 4.5e-10, solv.sig, np.zeros(2), solv.q,
                            solv.rho, solv.xvv)
    ecoul = oracle.atom_sums(1, grid, box, solv_q, sol_q, g)[:, 0] / 4184.0
    ad2 = oracle.atom_sums(2, grid, box, solv, sol, g, adswitch=0)
    assert np.allclose(ad2, ecoul, rtol=1e-13)


def test_cutoff_excludes_the_atom_site(oracle):
    """An atom on a grid point: the r = 0 point (and every r < sigma/2) is
    left out, so the sums stay finite."""
    solv, sol, grid, box, g = _setup(oracle)
    sol0 = oracle.Solute(sol.sig, sol.eps, sol.q, np.array([[0.0, 0.0, 0.0], [1.0, 0.5, 0.5]]))
    for kind in (0, 1, 2):
        assert np.all(np.isfinite(oracle.atom_sums(kind, grid, box, solv, sol0, g)))


def test_euv_single_point_closed_form(oracle):
    """g = 1 at two grid points and 0 elsewhere, one point inside sigma_uv / 2
    of the atom (left out) and one outside: euv reduces to dv rho_v E_uv(r),
    the pair energy of one solvent site at distance r, E_uv = 4 eps_uv
    [(sigma_uv / r)^12 - (sigma_uv / r)^6] + q_u q_v e^2 / (4 pi eps0 r) per
    mole, with the Lorentz-Berthelot pair parameters and the Coulomb constant
    from scipy's CODATA values (the oracle's Hartree x Bohr x N_A, from the
    reference's physical.h, agrees with them to 2e-7).  A wrong volume
    element, density weight, unit or mixing rule changes the result by far
    more than the tolerance."""
    from scipy import constants as C

    ke = C.e ** 2 / (4.0 * math.pi * C.epsilon_0) * C.N_A * 1e10  # J A / mol
    solv, sol, _, _, _ = _setup(oracle)
    sol1 = oracle.Solute(sol.sig[:1], sol.eps[:1], sol.q[:1], np.array([[0.1, -0.2, 0.05]]))
    grid, box = (16, 16, 16), (8.0, 8.0, 8.0)
    nx, ny, nz = grid
    dv = 0.125
    N = nx * ny * nz
    p_in, p_out = (8, 8, 8), (12, 8, 5)  # (x, y, z) grid indices
    g = np.zeros((solv.natv, N))
    for x, y, z in (p_in, p_out):
        g[:, (z * ny + y) * nx + x] = 1.0
    got = oracle.atom_sums(1, grid, box, solv, sol1, g)[0]
    r = np.linalg.norm(np.array([(p_out[0] - 8) * 0.5, (p_out[1] - 8) * 0.5,
                                 (p_out[2] - 8) * 0.5]) - sol1.xyz[0])
    r_in = np.linalg.norm(sol1.xyz[0])
    exp_site = []
    for v in range(solv.natv):
        sg = 0.5 * (sol1.sig[0] + solv.sig[v])
        assert r_in < 0.5 * sg <= r
        eps = math.sqrt(sol1.eps[0] * solv.eps[v])
        e_lj = 4.0 * eps * ((sg / r) ** 12 - (sg / r) ** 6)
        e_c = ke * sol1.q[0] * solv.q[v] / r
        exp_site.append(dv * solv.rho[v] * (e_lj + e_c))
    assert got[1:] == pytest.approx(exp_site, rel=1e-6)
    assert got[0] == pytest.approx(sum(exp_site), rel=1e-6)
