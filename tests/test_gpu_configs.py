"""GPU parity on the benchmark workloads at their own sizes, against the FP64
oracle (oracle/rism_oracle.c, pinned in tests/test_oracle_*.py).

  cfg1   64^3 / 32 A, 24-atom random-walk solute (seed 1001), TIP3P water
         (tip3p_drism.xvk), KH: full solves with AN2 2 0.6 0.2 and with
         Anderson(10), and the first 40 steps of each
  cfg2   512^3 / 256 A, 12877-atom cluster (seed 1002), 0.2 M NaCl TIP3P
         (tip3p_nacl02_drism.xvk), KH: potentials on sampled rows, one step
         of the full grid
  cfg3   256^3 / 128 A, 3000-atom cluster (seed 1003), 0.1 M KCl TIP3P
         (tip3p_kcl01_drism.xvk), PSE-3: the same checks
  cfg4   random-walk solutes 0, 1, 2 and 28 of the batch (grids 60^3,
         80x72x80, 64x80x64, 96x96x72), KH: one step each
  cfg5   medium clusters 3 (144^3) and 0 (on a 180x192x160 grid, which
         covers the remaining lengths of the medium sizes), 0.2 M NaCl, HNC:
         one step each

At 512^3 and 256^3 the oracle's own potentials would take hours on the host,
so the oracle step there runs on the GPU's u, fr and fk (each checked
against the oracle on sampled rows first), with the oracle's own k-shells
and chi table.  Step-level gates are the FP64 rounding gates of
test_gpu_solver.py (1e-10 of the field's largest magnitude, 1e-11 on the
rms) or, where larger, 100x the oracle's own response to a one-ulp
perturbation of the initial t (a strongly charged 12877-atom cluster with
ion sites, whose chi reaches 275, amplifies rounding far beyond the small
cases: the measured rounding sensitivity is the scale); full solves are gated at the reproducibility band of a solve
converged to 1e-6 (trajectories that differ only in rounding separate near
convergence): 2 J/mol on the free energies, 1e-5 L/mol on the PMV.
"""
import os
import time

import numpy as np
import pytest

import gen_inputs as G

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[configs]", *a, flush=True)


def _solvent(oracle, golden, name):
    s = oracle.read_xvk(os.path.join(golden, name))
    d = {"natv": s.natv, "ntab": s.ntab, "dr": s.dr, "temper": s.temper, "xt": s.xt,
         "vsig": s.sig, "veps": s.eps, "vq": s.q, "vrho": s.rho, "xvv": s.xvv}
    return s, d


def _centred(oracle, sol):
    so = oracle.Solute(np.asarray(sol["sig"], float), np.asarray(sol["eps"], float),
                       np.asarray(sol["q"], float), np.asarray(sol["xyz"], float))
    so.xyz, so.shift = oracle.centering(so.xyz)
    return so


def _solute(so):
    return {"usig": so.sig, "ueps": so.eps, "uq": so.q, "uxyz": so.xyz}


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ---------------------------------------------------------------------------
# cfg1: the 64^3 case exactly as configured
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cfg1(oracle, golden):
    so = _centred(oracle, G.cfg1_solute())
    s, d = _solvent(oracle, golden, "tip3p_drism.xvk")
    return so, s, d, (64, 64, 64), (32.0, 32.0, 32.0)


@pytest.mark.parametrize("am", [0, 10])
def test_cfg1_trajectory(rism, oracle, gpu, cfg1, am):
    """The first 40 steps (2 Picard + 38 mixing) with AN2 (am = 0) and with
    Anderson(10): rms and t follow the oracle."""
    so, s, d, grid, box = cfg1
    o = oracle.OracleRISM(grid, box, s, so, "KH", count=2, m=0.6, mp=0.2, tol=1e-12).setup()
    ro = o.solve(maxstep=40) if am == 0 else o.solve_anderson_m(am, maxstep=40)
    g = rism.RISM3D(grid, box, d, _solute(so), anderson_m=am)
    g.initialize()
    st, n, rms = g.iterate(maxstep=40, tol=1e-12)
    t = g.fields(rism.FIELD_T).reshape(-1)
    g.close()
    _say("cfg1 am", am, "rms", rms, ro.rms[-1], "t", _rel(t, ro.t))
    assert (st, n) == (rism.STATUS_MAXSTEP, 40)
    assert abs(rms - ro.rms[-1]) < 1e-8 * ro.rms[-1]
    assert _rel(t, ro.t) < 1e-8


@pytest.mark.parametrize("am", [0, 10])
def test_cfg1_full_solve(rism, oracle, gpu, cfg1, am):
    """Converged to 1e-6: both converge, in step counts within 5 of each
    other, to the same thermodynamics within the reproducibility band."""
    so, s, d, grid, box = cfg1
    o = oracle.OracleRISM(grid, box, s, so, "KH", count=2, m=0.6, mp=0.2, tol=1e-6,
                          maxstep=3000).setup()
    ro = o.solve() if am == 0 else o.solve_anderson_m(am)
    to = o.thermo()
    g = rism.RISM3D(grid, box, d, _solute(so), anderson_m=am)
    th, tm = g.run(maxstep=3000, tol=1e-6)
    h = g.fields(rism.FIELD_H).reshape(-1)
    g.close()
    _say("cfg1 am", am, "steps", tm["steps"], ro.steps, "SFE_SC", th["SFE_SC"], to["SFE_SC"],
         "PMV", th["PMV"], to["PMV"], "h", _rel(h, ro.h))
    assert ro.status == 0 and tm["status"] == rism.STATUS_CONVERGED
    assert abs(tm["steps"] - ro.steps) <= 5
    for k in ("SFE_SC", "SFE_SC_HNC", "SFE_GF"):
        assert abs(th[k] - to[k]) < 2.0, (k, th[k], to[k])
    assert abs(th["PMV"] - to["PMV"]) < 1e-5
    assert _rel(h, ro.h) < 1e-4


# ---------------------------------------------------------------------------
# full-size single steps (cfg2, cfg3) and the batch members (cfg4, cfg5)
# ---------------------------------------------------------------------------
def _rows(grid):
    """(z0, nzs, y0, nys) blocks of rows for the potential check: through the
    solute at the centre, beside it, and at the far corner; all with
    z0 >= nz/2, where the stored half spectrum of fk lives."""
    nx, ny, nz = grid
    return [(nz // 2, 2, ny // 2 - 1, 2), (nz // 2 + nz // 16, 1, ny // 2 + ny // 10, 1),
            (nz - 1, 1, 0, 1)]


def _abs_sums(oracle, grid, box, s, so, z0, nzs, y0, nys):
    """Sums of the magnitudes of the terms of u_v (per site), fr and fk over
    the solute atoms, on the same block of rows: with 10^4 atoms the sums
    cancel, and rounding in a different summation order is bounded by
    FP64 epsilon times these, not times the result."""
    nx, ny, nz = grid
    dr = np.asarray(box, float) / np.array(grid)
    z, y, x = np.meshgrid(np.arange(z0, z0 + nzs), np.arange(y0, y0 + nys), np.arange(nx),
                          indexing="ij")
    pts = np.stack([((x - nx // 2) * dr[0]).ravel(), ((y - ny // 2) * dr[1]).ravel(),
                    ((z - nz // 2) * dr[2]).ravel()], 1)
    kf = 2.0 * np.pi / np.asarray(box, float)
    kk = (np.stack([(x - nx / 2.0 + 0.5).ravel(), (y - ny / 2.0 + 0.5).ravel(),
                    (z - nz / 2.0 + 0.5).ravel()], 1) * kf) ** 2
    k2 = kk.sum(1)
    beta = 1.0 / (oracle.AVOGADRO * oracle.BOLTZMANN * s.temper)
    ub = oracle.lib().or_ubeta(s.temper, 1.0)
    slj = np.zeros((s.natv, len(pts)))
    sq = np.zeros(len(pts))
    for a0 in range(0, so.natu, 512):
        a1 = min(so.natu, a0 + 512)
        r2 = ((pts[:, None, :] - so.xyz[None, a0:a1, :]) ** 2).sum(-1)
        r = np.sqrt(r2)
        with np.errstate(divide="ignore"):
            sq += np.where(r >= 1e-5, np.abs(so.q[a0:a1]) / r, np.abs(so.q[a0:a1])).sum(1)
        for v in range(s.natv):
            sg2 = (0.5 * (so.sig[a0:a1] + s.sig[v])) ** 2
            e4 = 4.0 * np.sqrt(so.eps[a0:a1] * s.eps[v])
            x2 = np.maximum(r2 / sg2, 4e-6)
            slj[v] += (e4 * (x2 ** -6 + x2 ** -3)).sum(1)
    shp = (nzs, nys, nx)
    su = [(beta * slj[v] + abs(s.q[v]) * ub * sq).reshape(shp) for v in range(s.natv)]
    sfk = (4.0 * np.pi * ub * np.abs(so.q).sum() / (k2 * (k2 + 1.0))).reshape(shp)
    return su, (ub * sq).reshape(shp), sfk


def _check_potentials_rows(rism, oracle, g, grid, box, s, so):
    """u, fr and fk of the GPU on sampled blocks of rows against the oracle,
    each point within 1e-12 of its sum of |terms| (condition-scaled: about
    100x the typical rounding of a 10^4-term FP64 sum, far below what a
    missing atom or a wrong factor moves)."""
    nx, ny, nz = grid
    fkh = g.fk_half()
    for z0, nzs, y0, nys in _rows(grid):
        u_o, fr_o, fk_o = oracle.potentials_rows(grid, box, s, so, z0, nzs, y0, nys)
        su, sfr, sfk = _abs_sums(oracle, grid, box, s, so, z0, nzs, y0, nys)
        for v in range(s.natv):
            ug = g.get(rism.FIELD_U, v).reshape(nz, ny, nx)[z0:z0 + nzs, y0:y0 + nys]
            err = np.abs(ug - u_o[v]) / su[v]
            _say("rows", z0, y0, "site", v, "max |du| / sum|terms|", err.max())
            assert err.max() < 1e-12, (z0, y0, v)
        frg = g.get(rism.FIELD_FR).reshape(nz, ny, nx)[z0:z0 + nzs, y0:y0 + nys]
        assert (np.abs(frg - fr_o) / sfr).max() < 1e-12, (z0, y0)
        m0 = z0 - nz // 2
        fkg = fkh[m0:m0 + nzs, y0:y0 + nys]
        assert (np.abs(fkg - fk_o) / sfk).max() < 1e-12, (z0, y0)
    return fkh


def _one_step_full(rism, oracle, grid, box, s, d, so, closure, check_rows):
    """One step of the HIP path from the initial t, against the oracle's step
    on the GPU's potentials (after checking those on sampled rows when
    check_rows) or on the oracle's own (otherwise)."""
    t0 = time.time()
    g = rism.RISM3D(grid, box, d, _solute(so), closure=oracle.closure_code(closure))
    g.initialize()
    _say(grid, closure, "context + potentials", f"{time.time() - t0:.1f} s")
    o = oracle.OracleRISM(grid, box, s, so, closure)
    if check_rows:
        fkh = _check_potentials_rows(rism, oracle, g, grid, box, s, so)
        _say(grid, "potential rows ok", f"{time.time() - t0:.1f} s")
        o.u = g.fields(rism.FIELD_U).reshape(-1)
        o.fr = g.get(rism.FIELD_FR)
        o.fk = oracle.full_spectrum(fkh).view(np.float64).reshape(-1)
        del fkh
        if grid[0] == grid[1] == grid[2] and box[0] == box[1] == box[2]:
            o.ga, o.indga = oracle.cubic_shells(grid, box)
        else:
            o.ga, o.indga = oracle.kshells(grid, box)
        o.chi = oracle.chi_shells(s, o.ga)
        o.t0 = (s.q[:, None] * o.fr[None, :]).reshape(-1)
    else:
        o.setup()
    rms_o, g_o, R_o, h_o = o.calculate(o.t0)
    del g_o
    _say(grid, "oracle step", f"{time.time() - t0:.1f} s", "rms", rms_o)
    # the step's own sensitivity to rounding: the oracle again from t0
    # perturbed by one unit in the last place (random sign per point); the
    # HIP path (other radix orders, other summation orders) may differ from
    # the oracle by up to 100x that, and by no less than the fixed FP64
    # rounding gates
    rng = np.random.default_rng(7)
    tp = o.t0 * (1.0 + 2.0 ** -52 * rng.choice([-1.0, 1.0], o.t0.size))
    rms_p, g_p, R_p, h_p = o.calculate(tp)
    del g_p, tp
    sens = {"rms": abs(rms_p - rms_o) / rms_o,
            rism.FIELD_R: np.abs(R_p - R_o).max() / np.abs(R_o).max(),
            rism.FIELD_H: np.abs(h_p - h_o).max() / np.abs(h_o).max()}
    del R_p, h_p
    _say(grid, "oracle sensitivity to 1 ulp of t", sens)
    rms_g = g.calculate()
    _say(grid, "gpu step rms", rms_g, "rel diff", abs(rms_g - rms_o) / rms_o)
    assert abs(rms_g - rms_o) < max(1e-11, 100.0 * sens["rms"]) * rms_o
    N = o.N
    for which, ref in ((rism.FIELD_R, R_o), (rism.FIELD_H, h_o)):
        scale = np.abs(ref).max()
        err = 0.0
        for v in range(s.natv):
            got = g.get(which, v)
            err = max(err, np.abs(got - ref[v * N:(v + 1) * N]).max())
        _say(grid, "field", which, "max err / max", err / scale)
        assert err < max(1e-10, 100.0 * sens[which]) * scale, which
    g.close()


@pytest.mark.timeout(900)
def test_cfg2_one_step_512(rism, oracle, gpu, golden):
    so = _centred(oracle, G.cfg2_solute())
    s, d = _solvent(oracle, golden, "tip3p_nacl02_drism.xvk")
    _one_step_full(rism, oracle, (512, 512, 512), (256.0, 256.0, 256.0), s, d, so, "KH", True)


@pytest.mark.timeout(600)
def test_cfg3_one_step_256(rism, oracle, gpu, golden):
    so = _centred(oracle, G.cfg3_solute())
    s, d = _solvent(oracle, golden, "tip3p_kcl01_drism.xvk")
    _one_step_full(rism, oracle, (256, 256, 256), (128.0, 128.0, 128.0), s, d, so, "PSE3",
                   True)


@pytest.mark.parametrize("i", [0, 1, 2, 28])
def test_cfg4_member_one_step(rism, oracle, gpu, golden, i):
    sol = G.batch_solute(i)
    grid, box = G.batch_grid(sol["xyz"])
    so = _centred(oracle, sol)
    s, d = _solvent(oracle, golden, "tip3p_drism.xvk")
    _one_step_full(rism, oracle, grid, box, s, d, so, "KH", False)


@pytest.mark.parametrize("i,grid", [(3, None), (0, (180, 192, 160))])
def test_cfg5_member_one_step(rism, oracle, gpu, golden, i, grid):
    sol = G.medium_solute(i)
    if grid is None:
        grid, box = G.batch_grid(sol["xyz"], sizes=G.MEDIUM_SIZES)
    else:
        box = tuple(G.BATCH_SPACING * n for n in grid)
    so = _centred(oracle, sol)
    s, d = _solvent(oracle, golden, "tip3p_nacl02_drism.xvk")
    _one_step_full(rism, oracle, grid, box, s, d, so, "HNC", False)


@pytest.mark.timeout(600)
def test_cfg4_batch_equals_sequential(rism, oracle, gpu, golden):
    """All 256 solutes of the batched workload: rism_run_batch gives the same
    results, bit for bit, as the same solves run one after another."""
    s, d = _solvent(oracle, golden, "tip3p_drism.xvk")
    cases = []
    for i in range(256):
        sol = G.batch_solute(i)
        grid, box = G.batch_grid(sol["xyz"])
        cases.append((grid, box, _solute(_centred(oracle, sol))))

    def make():
        return [rism.RISM3D(gr, bx, d, su) for gr, bx, su in cases]

    gs = make()
    out = rism.run_batch(gs)
    for g in gs:
        g.close()
    gs = make()
    seq = []
    for g in gs:
        th, tm = g.run()
        seq.append((th, tm["status"], tm["steps"]))
        g.close()
    conv = sum(1 for o in out if o[1] == rism.STATUS_CONVERGED)
    _say("cfg4 batch: converged", conv, "of", len(out))
    assert conv == len(out)
    for (tb, sb, nb), (ts, ss, ns) in zip(out, seq):
        assert (sb, nb) == (ss, ns)
        for k in ("SFE_SC", "SFE_GF", "PMV"):
            assert tb[k] == ts[k]
