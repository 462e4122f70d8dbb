"""The XRISM .xvk generator (3d-rism-cuda-mpi_amd/xrism1d.py; the reference
consumes .xvk files and has no generator).  Pinned by the reference's own
bundled solvent files: solving for the sites, temperature and grid in each
file's header must give the chi_vv(k) table stored in it (XRISM-KH), and its
compressibility; plus the ideal-gas limit and a round trip through the
library's .xvk reader."""
import os

import numpy as np
import pytest

import xrism1d as X


def _table(path, natv):
    with open(path) as f:
        lines = f.read().split("\n")
    num = int(lines[1].split()[1])
    data = " ".join(l for l in lines[4 + num + 1:] if l.strip() and not l.startswith("#"))
    return np.array(data.split(), dtype=float).reshape(-1, natv, natv)


@pytest.mark.parametrize("name,tol", [
    ("h2o_vv.xvk", 3e-7), ("ctip3p.xvk", 2e-5),
    pytest.param("kclaq_vv.xvk", 5e-7, marks=pytest.mark.slow),
    pytest.param("naclaq_vv.xvk", 5e-7, marks=pytest.mark.slow)])
def test_regenerates_bundled_xvk(golden, name, tol):
    src = os.path.join(golden, name)
    with open(src) as f:
        f.readline()
        natv = int(f.readline().split()[2])
        xt_ref = float(f.readline().split()[3])
    res = X.regenerate(src)
    ref = _table(src, natv)
    assert res.chi.shape == ref.shape
    assert np.abs(res.chi - ref).max() < tol
    assert res.xt == pytest.approx(xt_ref, rel=2e-5)


def test_ideal_gas_limit():
    """No interactions and a vanishing density: h -> 0, so chi = w (the
    intramolecular matrix, j0(k l) summed over equivalent sites)."""
    sites = X.water_tip3p(density=1e-6)
    for s in sites:
        s.eps, s.q, s.sig = 0.0, 0.0, 1e-3
    res = X.solve(sites, 300.0, dr=0.05, ntab=1024)
    k = res.k
    l_oh = 0.9572
    l_hh = np.linalg.norm(np.subtract(sites[1].xyz, sites[2].xyz))
    w_oh = np.sin(k * l_oh) / (k * l_oh)
    w_hh = np.sin(k * l_hh) / (k * l_hh)
    assert np.allclose(res.chi[:, 0, 0], 1.0, atol=1e-9)
    assert np.allclose(res.chi[:, 0, 1], 2 * w_oh, atol=1e-9)
    assert np.allclose(res.chi[:, 1, 0], w_oh, atol=1e-9)
    assert np.allclose(res.chi[:, 1, 1], 1.0 + w_hh, atol=1e-9)


def test_written_file_reads_back(rism, tmp_path):
    """write_xvk output parsed by the library's reader (rism_read_xvk)."""
    sites = X.water_tip3p(density=55.0)
    res = X.solve(sites, 298.15, dr=0.05, ntab=2048)
    path = tmp_path / "w.xvk"
    X.write_xvk(path, res, 298.15, 0.05)
    d = rism.read_xvk(str(path))
    assert d["natv"] == 2 and d["ntab"] == 2048
    assert d["temper"] == pytest.approx(298.15)
    assert d["xt"] == pytest.approx(res.xt, rel=1e-7)
    assert np.allclose(np.asarray(d["xvv"]).reshape(res.chi.shape), res.chi, rtol=1e-14, atol=0)
    assert np.allclose(d["vq"], [-0.834, 0.417])


def test_dielectric_constants(golden):
    """The k -> 0 charge sum rule: XRISM water has eps = 1 + 3y (y = 4 pi beta
    rho mu^2 / 9, the textbook XRISM result); DRISM reaches the requested
    dielectric constant."""
    sites, temper, dr, _ = X.read_header(os.path.join(golden, "h2o_vv.xvk"))
    res = X.solve(sites, temper, dr, 2048)
    rho = sites[0].density * X.AVOGADRO * 1e-27
    mu = 2 * 0.417 * 0.58588
    ub = X.HARTREE * X.BOHR / (X.BOLTZMANN * temper)
    y = 4 * np.pi * ub * rho * mu ** 2 / 9
    assert X.dielectric_constant(res, sites, temper) == pytest.approx(1 + 3 * y, rel=1e-4)
    for eps in (78.4, 40.0):
        r = X.solve(sites, temper, dr, 2048, dielectric=eps)
        assert X.dielectric_constant(r, sites, temper) == pytest.approx(eps, rel=2e-4)


@pytest.mark.slow
def test_drism_salt_converges():
    """DRISM with ions (0.2 M NaCl in TIP3P water, the salt of the 512^3
    case): converges, and chi_vv(k) tends to w at large k."""
    sites = X.water_tip3p(density=55.2) + [X.ion("Na", 2, 3, 2.439, 0.0874, 1.0, 0.2),
                                           X.ion("Cl", 3, 4, 4.478, 0.0356, -1.0, 0.2)]
    res = X.solve(sites, 298.15, 0.05, 4096, dielectric=78.4)
    assert res.rms < 1e-10
    assert np.allclose(res.chi[-1, 2, 2], 1.0, atol=1e-6)
