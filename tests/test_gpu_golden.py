"""End-to-end known answers on the GPU.

1. example_milc/mobley_3053621.xmu / .xmu0 (the reference's bundled outputs;
   KH, ctip3p.xvk, 64 A box, 256^3, tol 1e-6, AN2 2 0.6 0.2).  The FP64 oracle
   reproduces them to 0.4 J/mol on SFE_SC (tests/golden/oracle_mobley.json);
   the remaining gap is the 1e-6 convergence tolerance reached along a
   different rounding path (the golden came from a 4-rank run).  Gates:
   |dSFE| <= 1.5 J/mol, |dPMV| <= 1e-5 L/mol, Pressure to 1e-11 relative,
   |dCorrection| <= 5 J/mol.
2. The oracle's own results for the same runs: the HIP path must agree to
   1e-6 relative (same algorithm, FP64 on both sides).
3. The CLI (rism3d) on the same input writes a .xmu that parses to the same
   numbers.
"""
import json
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("q0", [False, True])
def test_mobley_known_answer(rism, oracle, gpu, golden, q0):
    g = rism.RISM3D.from_inp(os.path.join(golden, "3drism.dat"),
                             os.path.join(golden, "mobley_3053621"), q0=q0, solvent_dir=golden)
    g.initialize()
    st, n, rms = g.iterate(10000, 1e-6)
    th = g.thermo()
    g.close()
    assert st == rism.STATUS_CONVERGED
    ref = oracle.parse_xmu(os.path.join(golden, "mobley_3053621.xmu" + ("0" if q0 else "")))
    print(q0, n, {k: th[k] for k in ("SFE_SC", "SFE_SC_HNC", "SFE_GF", "PMV", "Correction_Term")})
    assert abs(th["SFE_SC"] - ref["SFE_SC"]) <= 1.5
    assert abs(th["SFE_SC_HNC"] - ref["SFE_SC_HNC"]) <= 1.5
    assert abs(th["SFE_GF"] - ref["SFE_GF"]) <= 1.5
    assert abs(th["PMV"] - ref["PMV"]) <= 1e-5
    assert abs(th["Pressure"] - ref["Pressure"]) <= 1e-11 * ref["Pressure"]
    assert abs(th["Correction_Term"] - ref["Correction_Term"]) <= 5.0
    vec = os.path.join(golden, "oracle_mobley.json")
    if os.path.exists(vec):
        o = json.load(open(vec))["xmu0" if q0 else "xmu"]
        assert abs(n - o["steps"]) <= 2
        for k in ("SFE_SC", "SFE_SC_HNC", "SFE_GF", "PMV", "Correction_Term"):
            assert abs(th[k] - o[k]) <= 1e-6 * abs(o[k]), (k, th[k], o[k])


def test_cli_mobley(rism, oracle, gpu, golden, tmp_path):
    exe = os.path.join(ROOT, "3d-rism-cuda-mpi_amd", "rism3d")
    for f in ("3drism.dat", "mobley_3053621", "ctip3p.xvk"):
        shutil.copy(os.path.join(golden, f), tmp_path / f)
    p = subprocess.run([exe, "-i", "3drism.dat", "-s", "mobley_3053621"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "relaxing 3D UV RISM:" in p.stdout
    assert " Step = 1 Reside = " in p.stdout
    got = oracle.parse_xmu(tmp_path / "mobley_3053621.xmu")
    ref = oracle.parse_xmu(os.path.join(golden, "mobley_3053621.xmu"))
    assert abs(got["SFE_SC"] - ref["SFE_SC"]) <= 1.5
    assert abs(got["PMV"] - ref["PMV"]) <= 1e-5
    text = open(tmp_path / "mobley_3053621.xmu").read().splitlines()
    assert text[0] == " $RESULTS" and text[-1] == " $END"
