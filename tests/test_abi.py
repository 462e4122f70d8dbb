"""The C-ABI library without a GPU: it loads, exports every function that
include/rism3d_c.h declares, and its host-only entry points (the .inp/.xvk
parsers, the .xmu writer, error codes) agree with the oracle's restatement of
read_input.cu / solvent_read.cc / output_xmu.cu.  No device call is made."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rism3d_c.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(rism_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(rism):
    L = rism.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(rism.EXPORTS)
    nm = subprocess.run(["nm", "-D", "--defined-only", rism.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}$", nm, re.M), n


def test_version(rism):
    assert rism.lib().rism_version().decode() == "20220301"


@pytest.mark.parametrize("name", ["h2o_vv.xvk", "naclaq_vv.xvk", "kclaq_vv.xvk", "ctip3p.xvk"])
def test_xvk_parser_matches_oracle(rism, oracle, golden, name):
    a = rism.read_xvk(os.path.join(golden, name))
    b = oracle.read_xvk(os.path.join(golden, name))
    assert (a["natv"], a["ntab"]) == (b.natv, b.ntab)
    assert a["dr"] == b.dr and a["temper"] == b.temper and a["xt"] == b.xt
    for k, v in (("vsig", b.sig), ("veps", b.eps), ("vq", b.q), ("vrho", b.rho)):
        assert np.array_equal(a[k], v), k
    assert np.array_equal(a["xvv"], b.xvv)


def test_xvk_site_merging(rism, golden):
    """solvent_read.cc:70-79: the two H lines of water merge into one site
    whose density is the sum (2 x 55.33296 M)."""
    a = rism.read_xvk(os.path.join(golden, "h2o_vv.xvk"))
    assert a["natv"] == 2
    assert np.isclose(a["vrho"][1] / a["vrho"][0], 2.0)
    assert np.isclose(a["vrho"][0], 55.33296 * 6.02214076e23 * 1e-27)
    n = rism.read_xvk(os.path.join(golden, "naclaq_vv.xvk"))
    assert n["natv"] == 4 and list(n["vq"]) == [-0.834, 0.417, 1.0, -1.0]


@pytest.mark.parametrize("args", [
    ("3drism.dat", "mobley_3053621", True, False),
    ("3drism.dat", "mobley_3053621", True, True),
    ("pegasus_test1.inp", None, True, False),
    ("pegasus_test1.inp", None, False, False),
    ("cygnus_test2.inp", None, True, False),
])
def test_inp_parser_matches_oracle(rism, oracle, golden, args):
    ctl, sol, center, q0 = args
    c = os.path.join(golden, ctl)
    s = os.path.join(golden, sol) if sol else None
    a = rism.read_input(c, s, center, q0, solvent_dir=golden)
    oc, os_ = oracle.read_inp(c, s, center, q0)
    assert a["closure"] == {"KH": 0, "HNC": 1}[oc.closure]
    assert a["fsolvent"] == oc.fsolvent and a["outlist"] == oc.outlist
    assert (a["count"], a["m"], a["mp"]) == (oc.count, oc.m, oc.mp)
    assert a["convergence"] == oc.convergence and a["maxstep"] == oc.maxstep
    assert np.array_equal(a["box"], oc.box) and np.array_equal(a["grid"], oc.grid)
    assert np.array_equal(a["shift"], os_.shift)
    for k, v in (("usig", os_.sig), ("ueps", os_.eps), ("uq", os_.q), ("uxyz", os_.xyz)):
        assert np.array_equal(a[k], v), k


def test_split_control_and_solute(rism, golden, tmp_path):
    """-i control -s solute reads the atoms from the second file."""
    a = rism.read_input(os.path.join(golden, "example_test1_c"),
                        os.path.join(golden, "example_test1_s"), solvent_dir=golden)
    b = rism.read_input(os.path.join(golden, "pegasus_test1.inp"), solvent_dir=golden)
    assert np.array_equal(a["uxyz"], b["uxyz"]) and np.array_equal(a["uq"], b["uq"])


def test_centering_known_answer(rism, golden):
    """solute.cu:36-38 on the water solute: x in [-0.75695, 0.75695] -> shift
    round(0) = 0; z in [0, 0.58588] -> round(-0.29294) = -0."""
    a = rism.read_input(os.path.join(golden, "pegasus_test1.inp"), solvent_dir=golden)
    assert np.array_equal(a["shift"], [0.0, 0.0, 0.0])
    b = rism.read_input(os.path.join(golden, "3drism.dat"), os.path.join(golden, "mobley_3053621"),
                        solvent_dir=golden)
    # x in [-0.08, 3.76] -> round(-1.92 + 0.08) = round(-1.84) = -2
    assert b["shift"][0] == -2.0


@pytest.mark.parametrize("text,code", [
    (" m 0 20210101\n KH\n", 3),
    (" m 0 20220301\n PY\n", 4),
    (" m 0 20220301\n KH\n h2o_vv.xvk\n 1e-6 100\n 2 0.6 0.2\n 8 8 8\n 16 16 16\n -3\n", 5),
    (" m 0 20220301\n KH\n missing.xvk\n 1e-6 100\n 2 0.6 0.2\n 8 8 8\n 16 16 16\n 0\n", 2),
])
def test_input_error_codes(rism, golden, tmp_path, text, code):
    p = tmp_path / "bad.inp"
    p.write_text(text)
    with pytest.raises(rism.RismError) as e:
        rism.read_input(str(p), solvent_dir=golden)
    assert e.value.code == code


def test_xmu_writer_format(rism, oracle, golden, tmp_path):
    """output_xmu.cu:28-79 layout, checked against the bundled .xmu totals."""
    ref = oracle.parse_xmu(os.path.join(golden, "mobley_3053621.xmu"))
    th = rism.Thermo()
    th.natv = 2
    th.sfe_sc, th.sfe_sc_hnc, th.sfe_gf = ref["SFE_SC"], ref["SFE_SC_HNC"], ref["SFE_GF"]
    th.pmv, th.pressure, th.correction = ref["PMV"], ref["Pressure"], ref["Correction_Term"]
    th.sfec_sc[0], th.sfec_sc[1] = 225232.23273, -160285.52378
    th.sfec_sc_hnc[0], th.sfec_sc_hnc[1] = 225232.23273, -160285.52378
    th.sfec_gf[0], th.sfec_gf[1] = 213595.33419, -175299.02902
    out = tmp_path / "x.xmu"
    rism.write_xmu(str(out), rism.KH, th)
    lines = out.read_text().splitlines()
    golden_lines = open(os.path.join(golden, "mobley_3053621.xmu")).read().splitlines()
    assert lines[0] == " $RESULTS" and lines[-1] == " $END"
    for want in ("SFE_SC= 64946.70895 !(J/mol)", "  SFEC_SC(0)= 225232.23273",
                 "SFE_SC_HNC= 77631.42548 !(J/mol)", "SFE_GF= 38296.30517 !(J/mol)",
                 "  SFEC_GF(1)= -175299.02902", "PMV= 0.07647 !(L/mol)",
                 "Pressure= 1164015832.87831 !(J/m^3)", "Correction_Term= -89014.34300 !(J/mol)"):
        assert want in lines, want
        assert want in golden_lines, want
    # HNC runs have no SC_HNC block
    rism.write_xmu(str(out), rism.HNC, th)
    assert not any(l.startswith("SFE_SC_HNC") for l in out.read_text().splitlines())


def test_cli_no_input(golden):
    exe = os.path.join(ROOT, "3d-rism-cuda-mpi_amd", "rism3d")
    p = subprocess.run([exe], capture_output=True, text=True)
    assert p.returncode == 1 and "No input file!" in p.stdout


def test_cli_old_version(golden, tmp_path):
    exe = os.path.join(ROOT, "3d-rism-cuda-mpi_amd", "rism3d")
    (tmp_path / "old.inp").write_text(" m 0 20200101\n KH\n")
    p = subprocess.run([exe, "old.inp"], cwd=tmp_path, capture_output=True, text=True)
    assert p.returncode == 1 and "old version" in p.stdout


def test_fft_lengths_outside_the_built_set_are_refused(rism):
    """Every even size up to 1024 that is not an instantiated FFT length
    (test_gpu_fft_lengths.LENGTHS, each checked against the oracle on the
    GPU) is refused with RISM_ERR_ARG before any device call."""
    from test_gpu_fft_lengths import LENGTHS
    L = rism.lib()
    box = np.array([4.0, 4.0, 8.0])
    for n in range(8, 1025, 2):
        if n in LENGTHS:
            continue
        buf = np.zeros(n * 8 * 16)
        rc = L.rism_k_fft_r2c(n, 8, 16, rism._ptr(box), rism._ptr(buf),
                              rism._ptr(np.zeros(n * 8 * 8 * 2)))
        assert rc == rism.RISM_ERR_ARG, n
