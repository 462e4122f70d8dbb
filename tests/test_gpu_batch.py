"""Batched solves (rism_solve_batch / rism_run_batch; no counterpart in the
reference, which runs one solute per process): many contexts of uneven grid
sizes advanced together, each on its own stream without a host round trip
per step; contexts of one grid, box, solvent and options share each
launch (one context per blockIdx.z).  The batch must give exactly the
results of separate solves done step by step (the same arithmetic per
context), including the step at which each one stops, whatever its outcome
(converged, diverged, step limit).
"""
import numpy as np
import pytest

import gen_inputs as G

pytestmark = pytest.mark.gpu


def _solvent(rism, golden):
    return rism.read_xvk(f"{golden}/h2o_vv.xvk")


def _solute(sol):
    xyz = sol["xyz"]
    mn, mx = xyz.min(0), xyz.max(0)
    shift = np.array([np.floor(-(mx[i] - mn[i]) / 2 - mn[i] + 0.5) for i in range(3)])
    return {"usig": sol["sig"], "ueps": sol["eps"], "uq": sol["q"], "uxyz": xyz + shift}


def _case(i):
    sol = G.batch_solute(i)
    grid, box = G.batch_grid(sol["xyz"], sizes=(48, 60, 64), pad=8.0)
    return grid, box, _solute(sol)


def _stepwise(rism, solv, grid, box, su, closure, maxstep, tol):
    """rism_run with per-kernel timing on: the step-by-step loop that reads the
    rms back after every step."""
    g = rism.RISM3D(grid, box, solv, su, closure=closure)
    g.kernel_timing(True)
    th, tm = g.run(maxstep=maxstep, tol=tol)
    t = g.fields(rism.FIELD_T)
    g.close()
    return th, tm, t


def test_sync_free_switch(rism, gpu, golden, monkeypatch):
    """RISM_SYNC_FREE=0 keeps rism_solve in the step-by-step loop (no kernel
    timing involved); its results equal the sync-free path's bit for bit."""
    solv = _solvent(rism, golden)
    grid, box, su = _case(1)
    out = []
    for flag in ("0", None):
        if flag is None:
            monkeypatch.delenv("RISM_SYNC_FREE", raising=False)
        else:
            monkeypatch.setenv("RISM_SYNC_FREE", flag)
        g = rism.RISM3D(grid, box, solv, su)
        th, tm = g.run(maxstep=1000, tol=1e-6)
        out.append((th, tm, g.fields(rism.FIELD_T)))
        g.close()
    (th0, tm0, t0), (th1, tm1, t1) = out
    assert tm0["status"] == tm1["status"] == rism.STATUS_CONVERGED
    assert tm0["steps"] == tm1["steps"]
    for k in ("SFE_SC", "SFE_GF", "PMV"):
        assert th0[k] == th1[k], k
    assert np.array_equal(t0, t1)


def test_batch_equals_stepwise(rism, gpu, golden):
    solv = _solvent(rism, golden)
    cases = [_case(i) for i in range(6)]
    print([c[0] for c in cases])
    ref = [_stepwise(rism, solv, g, b, su, rism.KH, 1000, 1e-6) for g, b, su in cases]
    gs = [rism.RISM3D(g, b, solv, su) for g, b, su in cases]
    out = rism.run_batch(gs, maxstep=1000, tol=1e-6, threads=3)
    for (th, st, steps), g, (thr, tmr, tr) in zip(out, gs, ref):
        assert st == tmr["status"] == rism.STATUS_CONVERGED
        assert steps == tmr["steps"]
        for k in ("SFE_SC", "SFE_GF", "PMV"):
            assert th[k] == thr[k], k
        assert np.array_equal(g.fields(rism.FIELD_T), tr)
        g.close()


def test_batch_mixed_outcomes(rism, gpu, golden):
    """A converging KH solve, an HNC solve that overflows (divergence) and one
    stopped by the step limit, in one batch."""
    solv = _solvent(rism, golden)
    g0, b0, su0 = _case(0)
    # the 24-atom solute whose HNC solve overflows exp(-u + t) (test_gpu_solver)
    strong = _solute(G.cfg1_solute())
    specs = [(g0, b0, su0, rism.KH, 1000), ((48, 48, 48), (24.0, 24.0, 24.0), strong, rism.HNC, 1000),
             (g0, b0, su0, rism.KH, 7)]
    ref = [_stepwise(rism, solv, g, b, su, cl, ms, 1e-6) for g, b, su, cl, ms in specs]
    assert ref[1][1]["status"] == rism.STATUS_DIVERGED
    assert ref[2][1]["status"] == rism.STATUS_MAXSTEP
    for (g, b, su, cl, ms), (thr, tmr, tr) in zip(specs, ref):
        h = rism.RISM3D(g, b, solv, su, closure=cl)
        (th, st, steps), = rism.run_batch([h], maxstep=ms, tol=1e-6)
        assert (st, steps) == (tmr["status"], tmr["steps"])
        if st != rism.STATUS_DIVERGED:
            assert th["SFE_SC"] == thr["SFE_SC"]
            assert np.array_equal(h.fields(rism.FIELD_T), tr)
        h.close()
    # all three together
    hs = [rism.RISM3D(g, b, solv, su, closure=cl) for g, b, su, cl, ms in specs]
    for h in hs:
        h.initialize()
    res = rism.solve_batch(hs[:2], maxstep=1000, tol=1e-6, threads=2)
    assert [tuple(r[:2]) for r in res] == [(t["status"], t["steps"]) for _, t, _ in ref[:2]]
    for h in hs:
        h.close()


def test_batch_with_host_accelerated_contexts(rism, gpu, golden):
    """Mixed-precision and Anderson(m) contexts, whose accelerators solve on
    the host every step, run in a batch beside FP64 AN2 contexts (the
    step-by-step loop on their worker thread): every member equals its
    separate solve bit for bit."""
    solv = _solvent(rism, golden)
    cases = [_case(i) for i in range(4)]
    kinds = [dict(), dict(precision=1), dict(anderson_m=4), dict(precision=1, anderson_m=3)]
    ref = []
    for (g, b, su), kw in zip(cases, kinds):
        h = rism.RISM3D(g, b, solv, su, **kw)
        th, tm = h.run(maxstep=1000, tol=1e-6)
        ref.append((th, tm, h.fields(rism.FIELD_T)))
        h.close()
    hs = [rism.RISM3D(g, b, solv, su, **kw) for (g, b, su), kw in zip(cases, kinds)]
    out = rism.run_batch(hs, maxstep=1000, tol=1e-6, threads=2)
    for (th, st, steps), h, (thr, tmr, tr), kw in zip(out, hs, ref, kinds):
        assert st == tmr["status"] == rism.STATUS_CONVERGED, kw
        assert steps == tmr["steps"], kw
        for k in ("SFE_SC", "SFE_GF", "PMV"):
            assert th[k] == thr[k], (kw, k)
        assert np.array_equal(h.fields(rism.FIELD_T), tr), kw
        h.close()


def test_grouped_members_stop_at_their_own_steps(rism, gpu, golden):
    """Contexts of one grid, box, solvent and closure share every launch of a
    batch (one context per blockIdx.z).  Neutral solutes of 6-16 atoms and
    the charged 24-atom solute, all HNC, in one group: their separate solves
    stop at different steps, converged or diverged, and each member stops
    where its separate step-by-step solve stops, with the same fields bit for
    bit; with a step limit of 7 every member stops where its separate solve
    does."""
    solv = _solvent(rism, golden)
    grid, box = (48, 48, 48), (24.0, 24.0, 24.0)
    sus = []
    for n, seed in ((6, 31), (9, 32), (12, 33), (16, 34)):
        sol = G.random_walk_solute(n, seed=seed, total_charge=0.0)
        su = _solute(sol)
        su["uq"] = np.zeros_like(su["uq"])
        sus.append(su)
    sus.append(_solute(G.cfg1_solute()))
    for maxstep in (1000, 7):
        ref = [_stepwise(rism, solv, grid, box, su, rism.HNC, maxstep, 1e-6) for su in sus]
        print(maxstep, [(t["status"], t["steps"]) for _, t, _ in ref])
        if maxstep == 1000:  # observed: converged at 147, 244, 249; diverged at 60, 2
            assert ref[-1][1]["status"] == rism.STATUS_DIVERGED
            assert sum(t["status"] == rism.STATUS_CONVERGED for _, t, _ in ref) >= 2
            assert len(set(t["steps"] for _, t, _ in ref)) >= 4
        hs = [rism.RISM3D(grid, box, solv, su, closure=rism.HNC) for su in sus]
        out = rism.run_batch(hs, maxstep=maxstep, tol=1e-6, threads=1)
        for (th, st, steps), h, (thr, tmr, tr) in zip(out, hs, ref):
            assert (st, steps) == (tmr["status"], tmr["steps"])
            if st != rism.STATUS_DIVERGED:
                for k in ("SFE_SC", "SFE_GF", "PMV"):
                    assert th[k] == thr[k], k
                assert np.array_equal(h.fields(rism.FIELD_T), tr)
            h.close()


def test_batch_rejects_repeated_context(rism, gpu, golden):
    solv = _solvent(rism, golden)
    g, b, su = _case(1)
    h = rism.RISM3D(g, b, solv, su)
    h.initialize()
    with pytest.raises(rism.RismError) as e:
        rism.solve_batch([h, h])
    assert e.value.code == rism.RISM_ERR_ARG
    h.close()
