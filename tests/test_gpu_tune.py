"""Every launch plan the autotuner can pick gives the reference step.

The forward z, residual z and y passes are built in up to three plans per
FFT length (rism_kzpass.h, rism_kypass.hip); a context creation times them
and keeps the fastest (rism_tune_report lists the choices).  For every
tuned length, as nz/2 on (16, 16, 2L) grids and as ny on (16, L, 16) grids
(0.5 A spacing), each offered plan is forced through RISM_TUNE_PLANS and one
step (a Picard step: forward z pass, y passes, residual pass) is checked
against the FP64 oracle: FP64 at 1e-10 of the largest value (rounding of
different radix orders), FP32 (the mixed solve's bulk) at 3e-5.
"""
import os

import numpy as np
import pytest

from gen_inputs import random_walk_solute

pytestmark = pytest.mark.gpu

TUNED = [48, 60, 64, 72, 80, 90, 96, 128, 144, 160, 180, 192, 256, 512, 1024]
PASSES = {"zfwd_closure": 0, "zinv_resid": 1, "ypass": 2}


@pytest.fixture(scope="module")
def water(oracle, golden):
    s = oracle.read_xvk(f"{golden}/h2o_vv.xvk")
    d = {"natv": s.natv, "ntab": s.ntab, "dr": s.dr, "temper": s.temper, "xt": s.xt,
         "vsig": s.sig, "veps": s.eps, "vq": s.q, "vrho": s.rho, "xvv": s.xvv}
    sol = random_walk_solute(6, seed=21, total_charge=0.0)
    sol = oracle.Solute(sol["sig"], sol["eps"], sol["q"], sol["xyz"])
    sol.xyz, sol.shift = oracle.centering(sol.xyz)
    su = {"usig": sol.sig, "ueps": sol.eps, "uq": sol.q, "uxyz": sol.xyz}
    return s, d, sol, su


def _plans_env(pass_index, v):
    p = [0, 0, 0]
    p[pass_index] = v
    return ",".join(str(x) for x in p)


def _step(rism, grid, box, d, su, prec, plans):
    old = os.environ.get("RISM_TUNE_PLANS")
    if plans is None:
        os.environ.pop("RISM_TUNE_PLANS", None)
    else:
        os.environ["RISM_TUNE_PLANS"] = plans
    try:
        g = rism.RISM3D(grid, box, d, su, precision=prec)
    finally:
        if old is None:
            os.environ.pop("RISM_TUNE_PLANS", None)
        else:
            os.environ["RISM_TUNE_PLANS"] = old
    g.initialize()
    rms = g.calculate()
    R = g.fields(rism.FIELD_R).reshape(-1)
    h = g.fields(rism.FIELD_H).reshape(-1)
    g.close()
    return rms, R, h


@pytest.mark.parametrize("axis", ["z", "y"])
@pytest.mark.parametrize("L", TUNED)
def test_every_plan_matches_oracle(rism, oracle, gpu, water, L, axis):
    s, d, sol, su = water
    grid = (16, 16, 2 * L) if axis == "z" else (16, L, 16)
    box = tuple(0.5 * n for n in grid)
    passes = ("zfwd_closure", "zinv_resid") if axis == "z" else ("ypass",)
    o = oracle.OracleRISM(grid, box, s, sol, "KH").setup()
    rms_o, g_o, R_o, h_o = o.calculate(o.t0)
    for prec, tol in ((0, 1e-10), (1, 3e-5)):
        # an unforced context times the plans of this grid
        _step(rism, grid, box, d, su, prec, None)
        rep = [e for e in rism.tune_report()
               if e["grid"] == list(grid) and e["dtype"] == ("f32" if prec else "f64")
               and e["pass"] in passes]
        for e in rep:
            assert e["ms"][e["choice"]] is not None
            for v, ms in enumerate(e["ms"]):
                if ms is None:
                    continue
                rms, R, h = _step(rism, grid, box, d, su, prec,
                                  _plans_env(PASSES[e["pass"]], v))
                er = np.abs(R - R_o).max() / np.abs(R_o).max()
                eh = np.abs(h - h_o).max() / np.abs(h_o).max()
                print(grid, prec, e["pass"], "plan", v, f"{ms:.4f} ms", "R", er, "h", eh)
                assert abs(rms - rms_o) < tol * rms_o
                assert er < tol and eh < tol, (e["pass"], v)
