"""GPU parity of the solver stages against the FP64 oracle, on seeded inputs at
sizes the oracle finishes in seconds.

  potentials  cal_LJ.cu / cal_Coulomb.cu / cal_potential.cu   rel 1e-12
  one step    calculate.cu + cal_rms.cu                       rel 1e-10
  AN2 path    anderson.cu (t after k steps, leg solution)     rel 1e-9
  full solve  iterate.cu + output_xmu quantities              rel 1e-7

All arithmetic is FP64 on both sides; the tolerances cover summation-order
rounding only.
"""
import numpy as np
import pytest

from gen_inputs import random_walk_solute

pytestmark = pytest.mark.gpu


def _solvent(oracle, golden, name="h2o_vv.xvk"):
    s = oracle.read_xvk(f"{golden}/{name}")
    d = {"natv": s.natv, "ntab": s.ntab, "dr": s.dr, "temper": s.temper, "xt": s.xt,
         "vsig": s.sig, "veps": s.eps, "vq": s.q, "vrho": s.rho, "xvv": s.xvv}
    return s, d


def _solute(sol):
    return {"usig": sol.sig, "ueps": sol.eps, "uq": sol.q, "uxyz": sol.xyz}


@pytest.fixture(scope="module")
def case24(oracle, golden):
    """cfg1-like: 24-atom random-walk solute in water, 48^3 grid in 24 A."""
    sol = random_walk_solute(24, seed=1001, total_charge=0.0)
    sol = oracle.Solute(sol["sig"], sol["eps"], sol["q"], sol["xyz"])
    sol.xyz, sol.shift = oracle.centering(sol.xyz)
    s, d = _solvent(oracle, golden)
    return sol, s, d, (48, 48, 48), (24.0, 24.0, 24.0)


def test_potentials(rism, oracle, gpu, case24):
    sol, s, d, grid, box = case24
    u, fr, fk = oracle.potentials(grid, box, s, sol)
    g = rism.RISM3D(grid, box, d, _solute(sol))
    g.initialize()
    ug = g.fields(rism.FIELD_U).reshape(-1)
    frg = g.get(rism.FIELD_FR)
    # the LJ core reaches ~1e13 at the s^2 floor: compare relative per point
    assert np.allclose(ug, u, rtol=1e-12, atol=1e-12)
    assert np.allclose(frg, fr, rtol=1e-12, atol=1e-13 * np.abs(fr).max())
    nx, ny, nz = grid
    fkc = fk.view(np.complex128).reshape(nz, ny, nx)[nz // 2:]
    fkg = g.fk_half()
    assert np.abs(fkg - fkc).max() < 1e-11 * np.abs(fkc).max()
    # initial t = q_v fr (initialize_tuv.cu:22)
    t0 = g.fields(rism.FIELD_T)
    assert np.allclose(t0, s.q[:, None] * fr[None, :], rtol=1e-12, atol=1e-13 * np.abs(fr).max())
    g.close()


@pytest.mark.parametrize("closure", ["KH", "HNC", "PSE3"])
def test_one_step(rism, oracle, gpu, case24, closure):
    sol, s, d, grid, box = case24
    o = oracle.OracleRISM(grid, box, s, sol, closure).setup()
    rms_o, g_o, R_o, h_o = o.calculate(o.t0)
    g = rism.RISM3D(grid, box, d, _solute(sol), closure=oracle.closure_code(closure))
    g.initialize()
    rms_g = g.calculate()
    assert abs(rms_g - rms_o) < 1e-11 * rms_o
    for which, ref in ((rism.FIELD_R, R_o), (rism.FIELD_G, g_o), (rism.FIELD_H, h_o)):
        got = g.fields(which).reshape(-1)
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() < 1e-10 * scale, which
    g.close()


def test_an2_trajectory(rism, oracle, gpu, case24):
    """t after 8 steps (2 Picard + 6 modified-Anderson) and the last leg()
    solution follow the oracle."""
    sol, s, d, grid, box = case24
    o = oracle.OracleRISM(grid, box, s, sol, "KH", count=2, m=0.6, mp=0.2, tol=1e-12).setup()
    ro = o.solve(maxstep=8)
    g = rism.RISM3D(grid, box, d, _solute(sol), count=2, m=0.6, mp=0.2)
    g.initialize()
    st, n, rms = g.iterate(maxstep=8, tol=1e-12)
    assert (st, n) == (rism.STATUS_MAXSTEP, 8)
    assert abs(rms - ro.rms[-1]) < 1e-9 * ro.rms[-1]
    t = g.fields(rism.FIELD_T).reshape(-1)
    assert np.abs(t - ro.t).max() < 1e-9 * np.abs(ro.t).max()
    x, sv, cnt, b = g.an2_state()
    assert cnt == 0 and b == 0
    g.close()


@pytest.mark.parametrize("closure,qscale", [("KH", 1.0), ("KH", 0.0), ("HNC", 0.0),
                                           ("PSE3", 0.0), ("PSE2", 0.3)])
def test_full_solve_vs_oracle(rism, oracle, gpu, case24, closure, qscale):
    """KH with the full charges; HNC (whose exp(-u + t) overflows for this
    strongly charged solute, see test_divergence) and PSE-n (whose series
    also runs away in the first steps at full charge) with scaled charges."""
    sol, s, d, grid, box = case24
    sol = oracle.Solute(sol.sig, sol.eps, sol.q * qscale, sol.xyz)
    o = oracle.OracleRISM(grid, box, s, sol, closure, tol=1e-6, maxstep=400).setup()
    ro = o.solve()
    to = o.thermo()
    g = rism.RISM3D(grid, box, d, _solute(sol), closure=oracle.closure_code(closure))
    g.initialize()
    st, n, rms = g.iterate(maxstep=400, tol=1e-6)
    print(closure, qscale, "steps", n, ro.steps)
    tg = g.thermo()
    assert st == ro.status == rism.STATUS_CONVERGED
    assert n == ro.steps
    for k in ("SFE_SC", "SFE_GF", "PMV", "Pressure", "Correction_Term"):
        assert abs(tg[k] - to[k]) <= 1e-7 * max(1.0, abs(to[k])), (k, tg[k], to[k])
    if closure == "KH":
        assert abs(tg["SFE_SC_HNC"] - to["SFE_SC_HNC"]) <= 1e-7 * abs(to["SFE_SC_HNC"])
    h = g.fields(rism.FIELD_H).reshape(-1)
    assert np.abs(h - ro.h).max() < 1e-7 * np.abs(ro.h).max()
    g.close()


def test_divergence(rism, oracle, gpu, case24):
    """HNC with the full charges of the 24-atom solute overflows exp(-u + t):
    both sides stop at the same step with a non-finite rms and report
    divergence (iterate.cu:50-54, 73-76)."""
    sol, s, d, grid, box = case24
    o = oracle.OracleRISM(grid, box, s, sol, "HNC", tol=1e-6, maxstep=50).setup()
    ro = o.solve()
    g = rism.RISM3D(grid, box, d, _solute(sol), closure=rism.HNC)
    g.initialize()
    st, n, rms = g.iterate(maxstep=50, tol=1e-6)
    g.close()
    assert ro.status == 1 and st == rism.STATUS_DIVERGED
    assert n == ro.steps
    assert not np.isfinite(rms)


def test_maxstep_limit(rism, oracle, gpu, case24):
    """maxstep reached: status MAXSTEP, the thermodynamics still evaluate
    (h of the last step, t after its AN2 update, as output.cu reads them)."""
    sol, s, d, grid, box = case24
    o = oracle.OracleRISM(grid, box, s, sol, "KH", tol=1e-6, maxstep=5).setup()
    ro = o.solve()
    to = o.thermo()
    g = rism.RISM3D(grid, box, d, _solute(sol))
    g.initialize()
    st, n, rms = g.iterate(maxstep=5, tol=1e-6)
    tg = g.thermo()
    g.close()
    assert ro.status == 2 and st == rism.STATUS_MAXSTEP and n == 5
    for k in ("SFE_SC", "SFE_GF", "PMV"):
        assert abs(tg[k] - to[k]) <= 1e-8 * max(1.0, abs(to[k])), (k, tg[k], to[k])


def test_lambda_switches(rism, oracle, gpu, case24):
    """-l (LJ sigma scaling, Coulomb off) and -e (Coulomb scaling) potentials."""
    sol, s, d, grid, box = case24
    for ad, lam in ((1, 0.7), (2, 0.4)):
        u, fr, fk = oracle.potentials(grid, box, s, sol, lam=lam, adswitch=ad)
        g = rism.RISM3D(grid, box, d, _solute(sol), ad_switch=ad, lam=lam)
        g.initialize()
        assert np.allclose(g.fields(rism.FIELD_U).reshape(-1), u, rtol=1e-12, atol=1e-12)
        assert np.allclose(g.get(rism.FIELD_FR), fr, rtol=1e-12,
                           atol=1e-13 * max(np.abs(fr).max(), 1.0))
        g.close()


@pytest.mark.parametrize("solvent", ["h2o_vv.xvk", "tip3p_nacl02_drism.xvk"])
@pytest.mark.parametrize("ad,lam", [(0, 1.0), (1, 0.8)])
def test_atoms_on_grid_points(rism, oracle, gpu, golden, solvent, ad, lam):
    """Atoms exactly on grid points: the s^2 >= 4e-6 floor (cal_LJ.cu:11-12)
    and the r < 1e-5 Coulomb guard bind there, so the potential kernel's
    clamped pair path runs, for 2 and 4 sites, with Coulomb (ad 0) and
    without it (ad 1: -l).  u reaches ~1e13 at those points."""
    s, d = _solvent(oracle, golden, solvent)
    grid, box = (32, 32, 32), (16.0, 16.0, 16.0)  # dr = 0.5: grid points at k/2 A
    xyz = np.array([[0.0, 0.0, 0.0], [1.5, -1.0, 2.5], [-2.0, 2.0, 0.0], [1.23, -0.77, 2.1]])
    sol = oracle.Solute(np.array([3.40, 2.96, 3.25, 2.50]), np.array([0.086, 0.21, 0.17, 0.015]),
                        np.array([0.4, -0.5, -0.2, 0.3]), xyz)
    u, fr, fk = oracle.potentials(grid, box, s, sol, lam=lam, adswitch=ad)
    assert np.abs(u).max() > 1e12  # the floor was reached
    g = rism.RISM3D(grid, box, d, _solute(sol), ad_switch=ad, lam=lam)
    g.initialize()
    assert np.allclose(g.fields(rism.FIELD_U).reshape(-1), u, rtol=1e-12, atol=1e-12)
    assert np.allclose(g.get(rism.FIELD_FR), fr, rtol=1e-12,
                       atol=1e-13 * max(np.abs(fr).max(), 1.0))
    g.close()


def test_krange_error(rism, oracle, gpu, golden):
    """|k|max beyond the chi table -> RISM_ERR_KRANGE (solvent_spline.cu:17-20)."""
    s, d = _solvent(oracle, golden)
    sol = {"usig": np.array([3.0]), "ueps": np.array([100.0]), "uq": np.array([0.0]),
           "uxyz": np.zeros((1, 3))}
    with pytest.raises(rism.RismError) as e:
        rism.RISM3D((64, 64, 64), (2.0, 2.0, 2.0), d, sol)
    assert e.value.code == rism.RISM_ERR_KRANGE


def test_anderson_m_trajectory(rism, oracle, gpu, case24):
    """Anderson(m) (rism_opts.anderson_m, no counterpart in the reference):
    t after 12 steps (2 Picard + 10 mixing steps, history filling up to m = 5
    and then rotating) follows the oracle's AndersonM, whose least squares is
    solved from the explicit difference vectors."""
    sol, s, d, grid, box = case24
    o = oracle.OracleRISM(grid, box, s, sol, "KH", count=2, m=0.6, mp=0.2, tol=1e-12).setup()
    ro = o.solve_anderson_m(5, maxstep=12)
    g = rism.RISM3D(grid, box, d, _solute(sol), count=2, m=0.6, mp=0.2, anderson_m=5)
    g.initialize()
    st, n, rms = g.iterate(maxstep=12, tol=1e-12)
    assert (st, n) == (rism.STATUS_MAXSTEP, 12)
    assert abs(rms - ro.rms[-1]) < 1e-9 * ro.rms[-1]
    t = g.fields(rism.FIELD_T).reshape(-1)
    assert np.abs(t - ro.t).max() < 1e-9 * np.abs(ro.t).max()
    g.close()


def test_anderson_m16_tiny_tile(rism, oracle, gpu, golden):
    """Anderson(16) on a 16^3 grid: the FP64 residual tile of nz = 16 has
    16 threads, fewer than the 17 sums per workgroup (R.R and 16 history
    dots), and every column must still be written.  20 steps (2 Picard, 18
    mixing, the history full from step 18) follow the oracle's AndersonM(16)
    to 1e-6 (its least squares is solved from the difference vectors, ours
    from their normal equations, whose conditioning at 16 vectors is the
    square of theirs); a missing column would put garbage into the m x m
    solve."""
    sol = random_walk_solute(6, seed=11, total_charge=0.0)
    sol = oracle.Solute(sol["sig"], sol["eps"], sol["q"], sol["xyz"])
    sol.xyz, sol.shift = oracle.centering(sol.xyz)
    s, d = _solvent(oracle, golden)
    grid, box = (16, 16, 16), (12.0, 12.0, 12.0)
    o = oracle.OracleRISM(grid, box, s, sol, "KH", count=2, m=0.6, mp=0.2, tol=1e-14).setup()
    ro = o.solve_anderson_m(16, maxstep=20)
    g = rism.RISM3D(grid, box, d, _solute(sol), anderson_m=16)
    g.initialize()
    st, n, rms = g.iterate(maxstep=20, tol=1e-14)
    t = g.fields(rism.FIELD_T).reshape(-1)
    g.close()
    print("rms", rms, ro.rms[-1], "t", np.abs(t - ro.t).max() / np.abs(ro.t).max())
    assert (st, n) == (rism.STATUS_MAXSTEP, 20)
    assert abs(rms - ro.rms[-1]) < 1e-6 * ro.rms[0]
    assert np.abs(t - ro.t).max() < 1e-6 * np.abs(ro.t).max()


def test_anderson_m_same_fixed_point_as_an2(rism, oracle, gpu, case24):
    """Anderson(10) and the reference's AN2 reach the same solution: solved
    to rms 1e-9, the .xmu quantities agree to 1e-6 relative."""
    sol, s, d, grid, box = case24
    out = {}
    for am in (0, 10):
        g = rism.RISM3D(grid, box, d, _solute(sol), anderson_m=am)
        g.initialize()
        st, n, rms = g.iterate(maxstep=2000, tol=1e-9)
        assert st == rism.STATUS_CONVERGED, (am, n, rms)
        out[am] = (n, g.thermo())
        g.close()
    print("steps AN2", out[0][0], "Anderson(10)", out[10][0])
    for k in ("SFE_SC", "SFE_GF", "PMV"):
        a, b = out[0][1][k], out[10][1][k]
        assert abs(a - b) <= 1e-6 * max(1.0, abs(a)), (k, a, b)
    assert out[10][0] < out[0][0]


def test_anderson_m_rejects_bad_history(rism, oracle, gpu, case24):
    sol, s, d, grid, box = case24
    with pytest.raises(rism.RismError) as e:
        rism.RISM3D(grid, box, d, _solute(sol), anderson_m=17)
    assert e.value.code == rism.RISM_ERR_ARG


@pytest.mark.parametrize("ad,lam", [(0, 1.0), (1, 0.8), (2, 0.6)])
def test_analysis_outputs(rism, oracle, gpu, case24, ad, lam):
    """.gra / .euv / ad (cal_grad, cal_euv, cal_ad1, cal_ad2) against the
    oracle's brute-force sums over the same g (the solved g of the context)."""
    sol, s, d, grid, box = case24
    g = rism.RISM3D(grid, box, d, _solute(sol), ad_switch=ad, lam=lam)
    g.initialize()
    g.iterate(maxstep=30, tol=1e-6)
    gg = g.fields(rism.FIELD_G).reshape(s.natv, -1)
    for kind in (rism.ANALYSIS_GRAD, rism.ANALYSIS_EUV, rism.ANALYSIS_AD):
        got = g.analysis(kind)
        ref = oracle.atom_sums(kind, grid, box, s, sol, gg, lam=lam, adswitch=ad)
        scale = np.abs(ref).max()
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() < 1e-11 * scale, (kind, np.abs(got - ref).max(), scale)
    g.close()


def _steps(g, n):
    """n reference loop steps (iterate.cu:48-72): calculate, then the AN2 update."""
    out = []
    for _ in range(n):
        out.append(g.calculate())
        g.anderson()
    return out


def test_tuv_checkpoint_restart(rism, oracle, gpu, case24, tmp_path):
    """.tuv checkpoint (write_tuv.cu:19-28, read_tuv.cu:6-29, restored t
    uploaded): the file holds the header and t at the stream's default 6
    significant digits; a context restarted from it runs the same residual
    sequence bit for bit as one given those values directly, and both reach
    the fixed point of an uninterrupted solve."""
    sol, s, d, grid, box = case24
    su = _solute(sol)
    g = rism.RISM3D(grid, box, d, su)
    g.initialize()
    _steps(g, 30)
    path = tmp_path / "case.tuv"
    g.write_tuv(path)
    t_mem = g.fields(rism.FIELD_T)
    g.close()
    tok = path.read_text().split()
    assert [int(x) for x in tok[:4]] == [grid[0], grid[1], grid[2], d["natv"]]
    saved = np.array(tok[4:], dtype=float).reshape(d["natv"], -1)
    six = np.array([float(f"{x:g}") for x in t_mem.ravel()]).reshape(saved.shape)
    assert np.array_equal(saved, six)
    # restart from the file against the same values uploaded directly
    a = rism.RISM3D(grid, box, d, su)
    a.initialize()
    a.read_tuv(path)
    assert np.array_equal(a.fields(rism.FIELD_T), saved)
    b = rism.RISM3D(grid, box, d, su)
    b.initialize()
    for v in range(d["natv"]):
        b.set(rism.FIELD_T, v, saved[v])
    ra, rb = _steps(a, 12), _steps(b, 12)
    assert ra == rb
    b.close()
    st, n, rms = a.iterate(maxstep=2000, tol=1e-6)
    assert st == rism.STATUS_CONVERGED
    th_a = a.thermo()
    a.close()
    ref = rism.RISM3D(grid, box, d, su)
    th_r, tm_r = ref.run(maxstep=2000, tol=1e-6)
    ref.close()
    assert tm_r["status"] == rism.STATUS_CONVERGED
    # the same fixed point, to the tolerance band of a converged solve
    assert abs(th_a["SFE_SC"] - th_r["SFE_SC"]) < 2.0
    assert abs(th_a["PMV"] - th_r["PMV"]) < 1e-5
    # a mismatched grid is refused (read_tuv.cu:14-20)
    c = rism.RISM3D((48, 48, 40), (24.0, 24.0, 20.0), d, su)
    c.initialize()
    with pytest.raises(rism.RismError) as e:
        c.read_tuv(path)
    assert e.value.code == rism.RISM_ERR_ARG
    c.close()


def test_six_sites_one_step(rism, oracle, gpu, case24, golden):
    """A 6-site solvent (water and the 4-site NaCl table side by side, chi
    block-diagonal: a synthetic mixture that exercises the x pass's 8-site
    instantiation, shell table and spline) against the oracle, one step."""
    sol, s2, _, _, _ = case24
    s4 = oracle.read_xvk(f"{golden}/naclaq_vv.xvk")
    n = min(s2.ntab, s4.ntab)
    assert s2.dr == s4.dr
    x = np.zeros((n, 6, 6))
    x[:, :2, :2] = s2.xvv[:n]
    x[:, 2:, 2:] = s4.xvv[:n]
    cat = lambda a, b: np.concatenate([a, b])
    s6 = oracle.Solvent(6, n, s2.dr, s2.temper, s2.xt, cat(s2.sig, s4.sig), cat(s2.eps, s4.eps),
                        cat(s2.q, s4.q), cat(s2.rho, s4.rho), x)
    d6 = {"natv": 6, "ntab": n, "dr": s6.dr, "temper": s6.temper, "xt": s6.xt, "vsig": s6.sig,
          "veps": s6.eps, "vq": s6.q, "vrho": s6.rho, "xvv": s6.xvv}
    grid, box = (32, 32, 32), (16.0, 16.0, 16.0)
    o = oracle.OracleRISM(grid, box, s6, sol, "KH").setup()
    rms_o, g_o, R_o, h_o = o.calculate(o.t0)
    for spline in (False, True):
        g = rism.RISM3D(grid, box, d6, _solute(sol), chi_spline=spline)
        g.initialize()
        rms_g = g.calculate()
        assert abs(rms_g - rms_o) < 1e-11 * rms_o, spline
        for which, ref in ((rism.FIELD_R, R_o), (rism.FIELD_H, h_o)):
            got = g.fields(which).reshape(-1)
            assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max(), (which, spline)
        g.close()


def test_empty_solute(rism, oracle, gpu, case24):
    """No solute atoms: u = fr = fk = 0 and t = 0, so g = 1, h = 0 and the
    first residual is exactly zero; the solve stops at step 1 with a zero
    excess chemical potential."""
    _, _, d, grid, box = case24
    empty = {"usig": np.zeros(0), "ueps": np.zeros(0), "uq": np.zeros(0),
             "uxyz": np.zeros((0, 3))}
    g = rism.RISM3D(grid, box, d, empty)
    th, tm = g.run(maxstep=100, tol=1e-6)
    assert (tm["status"], tm["steps"], tm["rms"]) == (rism.STATUS_CONVERGED, 1, 0.0)
    assert not np.any(g.fields(rism.FIELD_H))
    assert th["SFE_SC"] == 0.0 and th["SFE_GF"] == 0.0
    g.close()
