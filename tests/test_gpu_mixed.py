"""Mixed-precision solve (rism_opts.precision = 1; no counterpart in the
reference): FP32 FFTs, closure, OZ and accelerator fields with FP64 sums,
then FP64 steps to the tolerance.  Checked against the FP64 oracle:

  one FP32 step      rms rel 1e-5; R, g, h within 3e-5 of max|ref|
  state transfer     FP64 -> FP32 -> FP64 keeps t to FP32 rounding
  converged solve    same .xmu quantities as the FP64 solve: |dSFE| well
                     inside the 0.05 kcal/mol (209.2 J/mol) gate
"""
import numpy as np
import pytest

from gen_inputs import random_walk_solute

pytestmark = pytest.mark.gpu

KCAL = 4184.0


@pytest.fixture(scope="module")
def case24(oracle, golden):
    """24-atom random-walk solute in water, 48^3 grid in 24 A."""
    sol = random_walk_solute(24, seed=1001, total_charge=0.0)
    sol = oracle.Solute(sol["sig"], sol["eps"], sol["q"], sol["xyz"])
    sol.xyz, sol.shift = oracle.centering(sol.xyz)
    s = oracle.read_xvk(f"{golden}/h2o_vv.xvk")
    d = {"natv": s.natv, "ntab": s.ntab, "dr": s.dr, "temper": s.temper, "xt": s.xt,
         "vsig": s.sig, "veps": s.eps, "vq": s.q, "vrho": s.rho, "xvv": s.xvv}
    su = {"usig": sol.sig, "ueps": sol.eps, "uq": sol.q, "uxyz": sol.xyz}
    return sol, s, d, su, (48, 48, 48), (24.0, 24.0, 24.0)


def test_fp32_one_step(rism, oracle, gpu, case24):
    sol, s, d, su, grid, box = case24
    o = oracle.OracleRISM(grid, box, s, sol, "KH").setup()
    rms_o, g_o, R_o, h_o = o.calculate(o.t0)
    g = rism.RISM3D(grid, box, d, su, precision=1)
    g.initialize()
    assert g.precision() == 1
    rms_g = g.calculate()
    assert abs(rms_g - rms_o) < 1e-5 * rms_o
    for which, ref in ((rism.FIELD_R, R_o), (rism.FIELD_G, g_o), (rism.FIELD_H, h_o)):
        got = g.fields(which).reshape(-1)
        scale = np.abs(ref).max()
        err = np.abs(got - ref).max()
        print(which, err / scale)
        assert err < 3e-5 * scale, which
    g.close()


def test_precision_transfer(rism, oracle, gpu, case24):
    """Moving the state FP64 -> FP32 -> FP64 rounds t once to FP32; a
    context without the FP32 level refuses the move."""
    sol, s, d, su, grid, box = case24
    g = rism.RISM3D(grid, box, d, su, precision=1)
    g.initialize()
    g.set_precision(0)
    g.iterate(maxstep=5, tol=1e-12)
    t64 = g.fields(rism.FIELD_T)
    g.set_precision(1)
    assert g.precision() == 1
    t32 = g.fields(rism.FIELD_T)
    assert np.allclose(t32, t64.astype(np.float32).astype(np.float64), rtol=0, atol=0)
    g.set_precision(0)
    assert np.array_equal(g.fields(rism.FIELD_T), t32)
    g.close()
    h = rism.RISM3D(grid, box, d, su)
    with pytest.raises(rism.RismError) as e:
        h.set_precision(1)
    assert e.value.code == rism.RISM_ERR_STATE
    h.close()


@pytest.mark.parametrize("closure,am,block", [("KH", 0, 0), ("HNC", 0, 0), ("KH", 5, 0),
                                              ("KH", 0, 3)])
def test_mixed_solve_matches_fp64(rism, oracle, gpu, case24, closure, am, block):
    """Mixed solve (FP32 phase, then FP32 difference-form blocks between FP64
    steps with the accelerator state carried across) against the FP64 solve:
    AN2 and Anderson(5), the default block and a block of 3 FP32 steps."""
    sol, s, d, su, grid, box = case24
    q = 1.0 if closure == "KH" else 0.0  # HNC diverges with the full charges
    su = dict(su, uq=su["uq"] * q)
    clos = {"KH": rism.KH, "HNC": rism.HNC}[closure]
    out = {}
    for prec in (0, 1):
        g = rism.RISM3D(grid, box, d, su, closure=clos, precision=prec, anderson_m=am,
                        mixed_block=block)
        th, tm = g.run(maxstep=2000, tol=1e-6)
        assert tm["status"] == rism.STATUS_CONVERGED, (prec, tm)
        assert g.precision() == 0
        out[prec] = (th, tm, g.fields(rism.FIELD_H))
        g.close()
    th64, tm64, h64 = out[0]
    th32, tm32, h32 = out[1]
    print(closure, am, block, "FP64 steps", tm64["steps"], "mixed steps", tm32["steps"], "of them FP32",
          tm32["steps_fp32"], "dSFE", th32["SFE_SC"] - th64["SFE_SC"])
    assert tm32["steps_fp32"] > tm32["steps"] // 2
    assert abs(th32["SFE_SC"] - th64["SFE_SC"]) < 0.01 * KCAL
    assert abs(th32["SFE_SC"] - th64["SFE_SC"]) < 1e-4 * abs(th64["SFE_SC"])
    assert abs(th32["PMV"] - th64["PMV"]) < 1e-4 * abs(th64["PMV"])
    assert np.sqrt(np.mean((h32 - h64) ** 2)) < 1e-4


def test_mixed_mobley_known_answer(rism, oracle, gpu, golden):
    """example_milc mobley_3053621 (256^3, KH): the mixed solve reproduces the
    bundled .xmu SFE_SC well inside the 0.05 kcal/mol gate."""
    g = rism.RISM3D.from_inp(f"{golden}/3drism.dat", f"{golden}/mobley_3053621",
                             solvent_dir=golden, precision=1)
    th, tm = g.run(maxstep=10000, tol=1e-6)
    g.close()
    ref = oracle.parse_xmu(f"{golden}/mobley_3053621.xmu")
    print("mixed mobley", th["SFE_SC"], ref["SFE_SC"], tm)
    assert tm["status"] == rism.STATUS_CONVERGED
    assert abs(th["SFE_SC"] - ref["SFE_SC"]) < 0.05 * KCAL
    assert abs(th["PMV"] - ref["PMV"]) < 2e-5


def test_mixed_fp32_overflow_restarts_in_fp64(rism, oracle, gpu, case24):
    """HNC at 0.3x the charges: exp(-u + t) overflows within two steps (above
    88 in FP32 and 709 in FP64).  When the plain FP32 phase overflows, the
    mixed solve restarts in FP64 from the initial t with a fresh accelerator:
    its outcome is the FP64 solve's, after the FP32 steps it has spent."""
    sol, s, d, su, grid, box = case24
    su = dict(su, uq=su["uq"] * 0.3)
    out = []
    for prec in (0, 1):
        g = rism.RISM3D(grid, box, d, su, closure=rism.HNC, precision=prec)
        th, tm = g.run(maxstep=500, tol=1e-6)
        out.append(tm)
        g.close()
    fp64, mixed = out
    print("FP64", fp64, "mixed", mixed)
    assert mixed["steps_fp32"] >= 1
    assert mixed["status"] == fp64["status"]
    assert mixed["steps"] == mixed["steps_fp32"] + fp64["steps"]


def test_mixed_restart_starts_from_the_t_set(rism, gpu, case24, capfd):
    """A t set through rism_set_field in a mixed context (here half the
    initial t) is also where the FP64 restart after an FP32 overflow starts:
    the first FP64 residual after the overflow equals the first residual of
    an FP64 context given the same t (with the set-up's t instead it would be
    ~300 against ~1e10)."""
    sol, s, d, su, grid, box = case24
    su = dict(su, uq=su["uq"] * 0.3)
    runs = {}
    x = None
    for prec in (0, 1):
        g = rism.RISM3D(grid, box, d, su, closure=rism.HNC, precision=prec, verbose=True)
        g.initialize()
        if x is None:
            x = 0.5 * g.fields(rism.FIELD_T).reshape(g.natv, -1)
        for v in range(g.natv):
            g.set(rism.FIELD_T, v, x[v])
        capfd.readouterr()
        g.iterate(maxstep=8, tol=1e-6)
        out = capfd.readouterr().out
        g.close()
        runs[prec] = [ln.split() for ln in out.splitlines() if ln.strip().startswith("Step =")]
        print(prec, runs[prec])
    first64 = runs[0][0][5]
    mixed = runs[1]
    k = next(i for i, w in enumerate(mixed) if "overflow:" in w)
    assert k + 1 < len(mixed)
    after = mixed[k + 1]
    assert after[3] == "Reside" and "(FP32)" not in after
    assert after[5] == first64, (after, first64)
