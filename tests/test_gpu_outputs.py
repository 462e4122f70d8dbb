"""Output files of one solve, written by the CLI (rism3d, host/rism3d_main.cpp)
for an input with outlist "mgdab", read back and checked against the same
solve through the C ABI:

* .guv (output_guv.cu:24-47): int32 natv, nx, ny, nz; f64 box[3]; f64 the
  centring shift[3]; then g of each site, x fastest, FP64 -- equal to
  rism_get_field(G) bit for bit (the CLI's step loop and rism_run run the
  same arithmetic);
* .gra (output_grad.cu:8-22): natu lines of the three gradient components,
  default stream formatting (6 significant digits);
* .euv (output_euv.cu:8-31): natu lines "total site_0 .. site_n-1", fixed
  (6 decimals);
* <base>-<adswitch>-<lambda> (output_ad.cu:8-22): natu lines.
"""
import os
import subprocess

import numpy as np
import pytest

import gen_inputs as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _read_guv(path):
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(16), dtype=np.int32)
        box = np.frombuffer(f.read(24), dtype=np.float64)
        shift = np.frombuffer(f.read(24), dtype=np.float64)
        natv, nx, ny, nz = (int(v) for v in hdr)
        g = np.frombuffer(f.read(), dtype=np.float64)
    return natv, (nx, ny, nz), box, shift, g.reshape(natv, -1) if g.size else g


@pytest.mark.parametrize("ad,lam", [(0, 1.0), (1, 0.8)])
def test_cli_output_files(rism, oracle, gpu, golden, tmp_path, ad, lam):
    exe = os.path.join(ROOT, "3d-rism-cuda-mpi_amd", "rism3d")
    os.symlink(os.path.join(golden, "h2o_vv.xvk"), tmp_path / "h2o_vv.xvk")
    sol = G.random_walk_solute(10, seed=77, total_charge=0.0)
    sol["xyz"] = sol["xyz"] + np.array([0.37, -0.21, 1.13])  # non-zero centring shift
    grid, box = (40, 32, 48), (20.0, 16.0, 24.0)
    G.write_inp(tmp_path / "walk.inp", sol, "h2o_vv.xvk", grid, box, outlist="mgdab")
    cmd = [exe] + ({1: ["-l", str(lam)], 2: ["-e", str(lam)]}.get(ad, [])) + ["walk.inp"]
    p = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr

    g = rism.RISM3D.from_inp(str(tmp_path / "walk.inp"), ad_switch=ad, lam=lam,
                             solvent_dir=str(tmp_path))
    th, tm = g.run(maxstep=10000, tol=1e-6)
    assert tm["status"] == rism.STATUS_CONVERGED
    gg = g.fields(rism.FIELD_G)
    natu = len(sol["sig"])

    natv, gr, bx, shift, gf = _read_guv(tmp_path / "walk.guv")
    assert (natv, gr) == (g.natv, tuple(grid))
    assert np.array_equal(bx, np.array(box))
    assert np.array_equal(shift, np.asarray(g.shift, dtype=float))
    assert np.any(shift != 0.0)
    assert np.array_equal(gf, gg.reshape(natv, -1))

    def rows(name):
        with open(tmp_path / name) as f:
            return np.array([[float(x) for x in line.split()] for line in f if line.strip()])

    gra = rows("walk.gra")
    ref = g.analysis(rism.ANALYSIS_GRAD)
    assert gra.shape == (natu, 3)
    assert np.allclose(gra, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())

    euv = rows("walk.euv")
    ref = g.analysis(rism.ANALYSIS_EUV)
    assert euv.shape == (natu, 1 + g.natv)
    assert np.allclose(euv, ref, rtol=0.0, atol=5.1e-7)
    assert np.allclose(euv[:, 0], euv[:, 1:].sum(1), rtol=0.0, atol=2e-6)

    adname = f"walk-{ad}-{lam:g}"
    adv = rows(adname).reshape(-1)
    ref = g.analysis(rism.ANALYSIS_AD)
    assert adv.shape == (natu,)
    assert np.allclose(adv, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())
    g.close()


def test_api_output_guv(rism, oracle, gpu, golden, tmp_path):
    """RISM3D.output (the 'm' and 'g' letters through the C ABI): the .xmu
    parses to the thermodynamics of the context and the .guv holds g."""
    g = rism.RISM3D.from_inp(os.path.join(golden, "pegasus_test1.inp"), solvent_dir=golden)
    g.initialize()
    g.iterate(maxstep=10, tol=1e-6)
    base = str(tmp_path / "t1")
    g.output(base)
    th = g.thermo()
    natv, gr, bx, shift, gf = _read_guv(base + ".guv")
    assert (natv, gr) == (g.natv, tuple(g.grid))
    assert np.array_equal(bx, np.array(g.box))
    assert np.array_equal(gf, g.fields(rism.FIELD_G).reshape(natv, -1))
    x = oracle.parse_xmu(base + ".xmu")
    for k in ("SFE_SC", "SFE_GF", "PMV", "Pressure"):
        assert abs(x[k] - th[k]) <= 1e-5 * max(1.0, abs(th[k])), k
    g.close()
