"""Pins the oracle's set-up stages on the CPU:
  * k-shells (initialize_g.cu): numpy unique of k^2;
  * natural spline (spline.cc / splint.cc): scipy CubicSpline(bc_type="natural");
  * chi at the shells (solvent_spline.cu): the same scipy spline on the
    bracketing knots of the .xvk table;
  * LJ / Coulomb / fk (cal_LJ.cu, cal_Coulomb.cu): brute-force numpy sums on
    1-3 atoms, a single-atom closed form, the r < 1e-5 and s^2 >= 4e-6 guards."""
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline


def test_kshells(oracle):
    grid, box = (8, 6, 10), (4.0, 3.5, 5.0)
    ga, ind = oracle.kshells(grid, box)
    nx, ny, nz = grid
    k = [2 * np.pi * (np.arange(n) - n / 2 + 0.5) / b for n, b in zip(grid, box)]
    kz, ky, kx = np.meshgrid(k[2], k[1], k[0], indexing="ij")
    k2 = (kx * kx + ky * ky + kz * kz).reshape(-1)
    u, inv = np.unique(k2, return_inverse=True)
    assert np.allclose(ga, np.sqrt(u), rtol=0, atol=0)
    assert np.array_equal(ind, inv)


@pytest.mark.parametrize("seed", [0, 1])
def test_spline_matches_scipy(oracle, seed):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.uniform(0.1, 1.0, 40)) if seed else np.linspace(0.3, 9.0, 40)
    y = np.sin(x) + rng.normal(0, 0.1, x.size)
    coe = np.zeros(3 * x.size)
    oracle.lib().or_spline_coef(x, y, x.size, coe)
    cs = CubicSpline(x, y, bc_type="natural")
    xs = np.concatenate([rng.uniform(x[0], x[-1], 200), x])
    got = np.array([oracle.lib().or_splint(x, y, coe, x.size, float(v)) for v in xs])
    assert np.allclose(got, cs(xs), rtol=1e-12, atol=1e-13)


def test_chi_shells_match_scipy(oracle, golden):
    s = oracle.read_xvk(os.path.join(golden, "naclaq_vv.xvk"))
    grid, box = (16, 16, 16), (12.0, 12.0, 12.0)
    ga, ind = oracle.kshells(grid, box)
    chi = oracle.chi_shells(s, ga).reshape(len(ga), s.natv, s.natv)
    tt = s.ttab
    lo = max(i for i in range(s.ntab) if tt[i] < ga[0]) if tt[0] < ga[0] else 0
    hi = min(i for i in range(s.ntab) if tt[i] > ga[-1])
    for v in range(s.natv):
        for w in range(s.natv):
            cs = CubicSpline(tt[lo:hi + 1], s.xvv[lo:hi + 1, v, w], bc_type="natural")
            assert np.allclose(chi[:, v, w], cs(ga), rtol=1e-11, atol=1e-12 * np.abs(
                s.xvv[lo:hi + 1, v, w]).max())


def test_chi_range_error(oracle, golden):
    s = oracle.read_xvk(os.path.join(golden, "h2o_vv.xvk"))
    ga, _ = oracle.kshells((64, 64, 64), (2.0, 2.0, 2.0))
    with pytest.raises(ValueError, match="insufficient maximal T"):
        oracle.chi_shells(s, ga)


def _grid_r(grid, box):
    r = [(np.arange(n) - n // 2) * (b / n) for n, b in zip(grid, box)]
    z, y, x = np.meshgrid(r[2], r[1], r[0], indexing="ij")
    return np.stack([x, y, z], -1).reshape(-1, 3)


def _solvent(oracle, golden):
    return oracle.read_xvk(os.path.join(golden, "h2o_vv.xvk"))


def test_lj_coulomb_fk_brute_force(oracle, golden):
    s = _solvent(oracle, golden)
    grid, box = (8, 8, 10), (4.0, 4.0, 5.0)
    sol = oracle.Solute(np.array([3.4, 2.5, 3.0]), np.array([300.0, 60.0, 800.0]),
                        np.array([-0.4, 0.3, 0.1]),
                        np.array([[0.0, 0.0, 0.0], [0.31, -0.77, 1.2], [1.5, 1.0, -0.9]]))
    u, fr, fk = oracle.potentials(grid, box, s, sol)
    R = _grid_r(grid, box)
    beta = 1.0 / (oracle.AVOGADRO * oracle.BOLTZMANN * s.temper)
    ub = oracle.HARTREE * oracle.BOHR / (oracle.BOLTZMANN * s.temper)
    d = np.linalg.norm(R[:, None, :] - sol.xyz[None, :, :], axis=2)
    de = np.where(d >= 1e-5, sol.q / np.where(d > 0, d, 1), 0.0).sum(1) * ub
    frb = np.where(d >= 1e-5, sol.q * (1 - np.exp(-d)) / np.where(d > 0, d, 1), sol.q).sum(1) * ub
    assert np.allclose(fr, frb, rtol=1e-13, atol=1e-14 * np.abs(frb).max())
    for v in range(s.natv):
        sig = (sol.sig + s.sig[v]) / 2
        eps = np.sqrt(sol.eps * s.eps[v])
        rs2 = np.maximum(d * d / sig ** 2, 4e-6)
        lj = (4 * eps * (rs2 ** -6 - rs2 ** -3)).sum(1) * beta
        assert np.allclose(u[v * R.shape[0]:(v + 1) * R.shape[0]], lj + s.q[v] * de, rtol=1e-12)
    k = [2 * np.pi * (np.arange(n) - n / 2 + 0.5) / b for n, b in zip(grid, box)]
    kz, ky, kx = np.meshgrid(k[2], k[1], k[0], indexing="ij")
    K = np.stack([kx, ky, kz], -1).reshape(-1, 3)
    k2 = (K * K).sum(1)
    fkb = (4 * np.pi * sol.q[None, :] * np.exp(-1j * K @ sol.xyz.T)).sum(1) / (k2 * (k2 + 1)) * ub
    assert np.allclose(fk.view(np.complex128), fkb, rtol=1e-12, atol=1e-14 * np.abs(fkb).max())


def test_single_atom_closed_form_and_guards(oracle, golden):
    s = _solvent(oracle, golden)
    grid, box = (8, 8, 8), (4.0, 4.0, 4.0)
    sig, eps, q = 3.0, 500.0, 0.5
    sol = oracle.Solute(np.array([sig]), np.array([eps]), np.array([q]), np.zeros((1, 3)))
    u, fr, fk = oracle.potentials(grid, box, s, sol)
    N = 512
    beta = 1.0 / (oracle.AVOGADRO * oracle.BOLTZMANN * s.temper)
    ub = oracle.HARTREE * oracle.BOHR / (oracle.BOLTZMANN * s.temper)
    # grid point (4,4,4) is r = 0: the atom sits on it
    i0 = 4 + 8 * 4 + 64 * 4
    assert fr[i0] == pytest.approx(q * ub, rel=1e-15)     # r < 1e-5: fr += q
    # r = 0: s^2 clamps to 4e-6 and de skips the atom
    for v in range(2):
        su = (sig + s.sig[v]) / 2
        eu = np.sqrt(eps * s.eps[v])
        irs6 = 1 / (4e-6) ** 3
        assert u[v * N + i0] == pytest.approx(beta * 4 * eu * irs6 * (irs6 - 1), rel=1e-14)
        # r = 1.5 A along x: closed form
        i1 = 7 + 8 * 4 + 64 * 4
        r = 1.5
        sr6 = (su / r) ** 6
        want = beta * 4 * eu * (sr6 * sr6 - sr6) + s.q[v] * ub * q / r
        assert u[v * N + i1] == pytest.approx(want, rel=1e-13)


def test_rows_equal_full_grid(oracle, golden):
    """potentials_rows (the oracle on a block of rows, for the full-size GPU
    checks) returns exactly the matching block of potentials() on the whole
    grid: the same loop over the same points."""
    s = _solvent(oracle, golden)
    grid, box = (12, 10, 16), (6.0, 5.0, 8.0)
    rng = np.random.default_rng(3)
    sol = oracle.Solute(rng.uniform(2.5, 3.5, 5), rng.uniform(60, 800, 5),
                        rng.normal(0, 0.4, 5), rng.uniform(-2, 2, (5, 3)))
    u, fr, fk = oracle.potentials(grid, box, s, sol)
    nx, ny, nz = grid
    z0, nzs, y0, nys = 5, 3, 2, 4
    ub, frb, fkb = oracle.potentials_rows(grid, box, s, sol, z0, nzs, y0, nys)
    u4 = u.reshape(s.natv, nz, ny, nx)
    assert np.array_equal(ub, u4[:, z0:z0 + nzs, y0:y0 + nys])
    assert np.array_equal(frb, fr.reshape(nz, ny, nx)[z0:z0 + nzs, y0:y0 + nys])
    fkc = fk.view(np.complex128).reshape(nz, ny, nx)
    assert np.array_equal(fkb, fkc[z0:z0 + nzs, y0:y0 + nys])


def test_cubic_shells_equal_sorted_shells(oracle):
    """cubic_shells (closed form, for 512^3 where sorting is slow) gives every
    point the same |k| as kshells (the sort of initialize_g.cu) to rounding."""
    for n, L in ((16, 8.0), (24, 11.0)):
        ga, ind = oracle.kshells((n, n, n), (L, L, L))
        ga2, ind2 = oracle.cubic_shells((n, n, n), (L, L, L))
        assert np.abs(ga[ind] - ga2[ind2]).max() <= 1e-15 * ga.max()


def test_full_spectrum_symmetry(oracle):
    """full_spectrum rebuilds the whole transform of a real field from its
    k_z > 0 half (F(-k) = conj F(k), k_{N-1-j} = -k_j)."""
    rng = np.random.default_rng(4)
    grid, box = (8, 12, 10), (4.0, 6.0, 5.0)
    f = rng.standard_normal((10, 12, 8))
    F = oracle.fft3d(f, box, -1)
    assert np.abs(oracle.full_spectrum(F[5:]) - F).max() <= 1e-14 * np.abs(F).max()
