"""Benchmark: seconds per converged 3D-RISM solve on one MI355X.

Headline workload (default): a 512^3 grid in a 256 A box; a synthetic
12877-atom protein-like cluster (uniform in a 30 A sphere, >= 1.0 A apart,
C/N/O/H + 1 % S, charges ~ N(0, 0.35 e) neutralised to +2 e, seed 1002);
0.2 M NaCl in TIP3P water (chi_vv from the project's 1D DRISM-KH at eps 78.4,
298.15 K, tests/golden/tip3p_nacl02_drism.xvk, 4 sites); KH closure; AN2
2 0.6 0.2; tolerance 1e-6; FP64 throughout.  One step = one full solve:
LJ/Coulomb potentials, the initial t, the fixed-point loop to convergence and
the thermodynamics, with the solute and the chi table already resident in
HBM when the clock starts.  The timed solves run the production path (no
per-kernel instrumentation).

The run is held to a wall-clock budget (--budget, default 450 s from process
start): warm-up and timed solves stop early when the next one would not fit,
and the line reports the counts actually run.  After the headline it adds,
while the budget lasts:

* roofline: the dominant hot-loop kernel's algorithmic bytes per launch over
  its average hipEvent-timed duration on the solver's stream, from a separate
  instrumented solve of the same input; traffic from two rocprofv3 --pmc
  child runs (FETCH_SIZE, WRITE_SIZE) scaled by the same counters on a plain
  copy of known size;
* mixed: the mixed-precision solve of the same input (s/solve, speed-up,
  |dmu|, dPMV and rms(dh) against the FP64 solve);
* batch: 256 small solutes solved as one batch against the same solves one
  after another;
* protein_test2: the reference's own 512^3 RBD-ACE2 input
  (example-cygnus/test2.inp, naclaq_vv.xvk), the case its published timings
  are quoted on;
* cpu_baseline: the FP64 oracle on the host cores (a bounded sample scaled to
  one solve, plus a full 64^3 solve);
* anderson10: the configuration's Anderson(10) accelerator, FP64 and mixed.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--budget S]
                  [--workload cfg2|cfg1|cfg3|mobley|test2|batch|batch-medium]
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import threading
import time

T_START = time.time()

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "3d-rism-cuda-mpi_amd"))

METRIC = ("s/solve at 512^3, batched solves/s; mixed vs FP64 at "
          "|Δμ|<0.05 kcal/mol")
HBM_PEAK_GBS = 8000.0
KCAL = 4184.0
# the reference's published time per solve of example-cygnus/test2.inp
# (README.md:25: 1 node, 4 GPUs, GPUDirect)
REF_PUBLISHED_TEST2_S = 413.9
# example-pegasus/test2.xmu:2,8,14 (an older build of the reference)
REF_TEST2_XMU = {"SFE_SC": 23324102.50402, "SFE_GF": 6329885.06176, "PMV": 69.89044}


def log(*a):
    print(f"[bench {time.time() - T_START:6.1f}s]", *a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# device plumbing (HIP runtime through ctypes; the solver owns its streams)
# ---------------------------------------------------------------------------
_hip = None


def hip():
    global _hip
    if _hip is None:
        import ctypes
        for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise SystemExit("bench.py needs the HIP runtime (libamdhip64.so)")
    return _hip


def device_count():
    import ctypes
    n = ctypes.c_int(0)
    if hip().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def sync():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def centred(sol):
    """solute.cu:18-48 centring (integer-A shift of the extent's midpoint)."""
    xyz = np.asarray(sol["xyz"], float)
    mn, mx = xyz.min(0), xyz.max(0)
    shift = np.array([np.floor(-(mx[i] - mn[i]) / 2 - mn[i] + 0.5) for i in range(3)])
    return dict(sol, xyz=xyz + shift)


def workload(name):
    import gen_inputs as G
    import __graft_entry__ as ge
    rism = ge.load_rism3d()
    if name == "cfg2":
        sol = centred(G.cfg2_solute())
        xvk = "tip3p_nacl02_drism.xvk"
        grid, box = (512, 512, 512), (256.0, 256.0, 256.0)
        desc = ("cfg2: 512^3 grid / 256 A box, synthetic 12877-atom cluster (R 30 A, +2 e, "
                "seed 1002), 0.2 M NaCl in TIP3P (chi from the project's 1D DRISM-KH, "
                "eps 78.4, 298.15 K, 4 sites), KH, AN2 2 0.6 0.2, tol 1e-6, FP64")
    elif name == "cfg1":
        sol = centred(G.cfg1_solute())
        xvk = "tip3p_drism.xvk"
        grid, box = (64, 64, 64), (32.0, 32.0, 32.0)
        desc = ("cfg1: 64^3 / 32 A, 24-atom random-walk solute (seed 1001), TIP3P water "
                "(chi from the project's 1D DRISM-KH, eps 78.4, 298.15 K), KH, AN2 2 0.6 0.2, "
                "tol 1e-6, FP64")
    elif name == "cfg3":
        sol = centred(G.cfg3_solute())
        xvk = "tip3p_kcl01_drism.xvk"
        grid, box = (256, 256, 256), (128.0, 128.0, 128.0)
        desc = ("cfg3: 256^3 / 128 A, synthetic 3000-atom cluster (R 19 A, 0 e, seed 1003), "
                "0.1 M KCl in TIP3P (chi from the project's 1D DRISM-KH, eps 78.4, 298.15 K, "
                "4 sites), tol 1e-6, FP64")
    elif name in ("mobley", "test2"):
        if name == "mobley":
            d = rism.read_input(os.path.join(GOLDEN, "3drism.dat"),
                                os.path.join(GOLDEN, "mobley_3053621"), solvent_dir=GOLDEN)
            desc = "example_milc mobley_3053621: 256^3 / 64 A, 12 atoms, ctip3p.xvk, KH, FP64"
        else:
            d = rism.read_input(os.path.join(GOLDEN, "cygnus_test2.inp"), solvent_dir=GOLDEN)
            desc = ("example-cygnus/test2.inp: RBD-ACE2, 12877 atoms, 512^3 / 256 A, "
                    "naclaq_vv.xvk, KH, AN2 2 0.6 0.2, tol 1e-6, FP64")
        sol = {"sig": d["usig"], "eps": d["ueps"], "q": d["uq"], "xyz": d["uxyz"]}
        xvk = os.path.basename(d["fsolvent"])
        grid, box = tuple(int(g) for g in d["grid"]), tuple(float(b) for b in d["box"])
    else:
        raise SystemExit(f"unknown workload {name}")
    solv = rism.read_xvk(os.path.join(GOLDEN, xvk))
    return rism, grid, box, solv, sol, xvk, desc


def solu_dict(sol):
    return {"usig": sol["sig"], "ueps": sol["eps"], "uq": sol["q"], "uxyz": sol["xyz"]}


def algorithmic_bytes(natv, N, hist=2, w=8):
    """Minimal HBM bytes per launch of each hot-loop kernel (w bytes per bulk
    word) with `hist` history residuals (2 for AN2, m for Anderson(m)):
    zfwd reads t, u (2n) and fr (1), writes the half spectrum (n);
    each y pass reads and writes the spectrum (2n);
    xpass_oz reads the spectrum (n) and the half fk (1), writes (n);
    zinv reads the spectrum, t, u (3n) and the history residuals (hist n),
    writes R (n);
    the update reads t, R and the history P_p (hist n), writes t', P_new (2n)."""
    ww = float(w) * N
    n = natv
    return {
        "zfwd_closure": (3 * n + 1) * ww,
        "ypass_fwd": 2 * n * ww,
        "xpass_oz": (2 * n + 1) * ww,
        "ypass_inv": 2 * n * ww,
        "zinv_resid": (4 + hist) * n * ww,
        "an2_update": (4 + hist) * n * ww,
    }


def fused_bytes(natv, N, m, w):
    """The fused lower bound of one iteration, each 3D transform counted as
    one HBM round trip: N [(12 + 3 m) n + 2] w bytes, m history pairs (2 for
    AN2), w bytes per bulk value."""
    return N * ((12 + 3 * m) * natv + 2) * w


class Budget:
    def __init__(self, total):
        self.end = T_START + total

    def left(self):
        return self.end - time.time()


# ---------------------------------------------------------------------------
# the phases of the default run
# ---------------------------------------------------------------------------
def timed_solves(g, warmup, steps, warm_cap_s, timed_cap_s, maxstep, label):
    """Warm-up solves (at least one, at most `warmup`, within warm_cap_s) then
    timed solves (at least one, at most `steps`, stopping when the next would
    end beyond timed_cap_s), each a full rism_run on the production path.
    Returns (warm-ups run, timed solves run, wall s, [(thermo, timing)])."""
    t0 = time.time()
    nw = 0
    per = None
    while nw < max(1, warmup):
        if nw >= 1 and time.time() - t0 + per > warm_cap_s:
            break
        ts = time.time()
        th, tm = g.run(maxstep=maxstep, tol=1e-6)
        per = time.time() - ts
        nw += 1
        log(f"{label} warm-up {nw}: {per:.2f} s, {tm['steps']} steps, status {tm['status']}")
    sync()
    out = []
    t1 = time.perf_counter()
    while len(out) < max(1, steps):
        el = time.perf_counter() - t1
        if out and el + el / len(out) > timed_cap_s:
            break
        th, tm = g.run(maxstep=maxstep, tol=1e-6)
        out.append((th, tm))
        log(f"{label} timed {len(out)}: {tm['steps']} steps, SFE_SC {th['SFE_SC']:.5f}")
    sync()
    wall = time.perf_counter() - t1
    return nw, len(out), wall, out


def kernel_profile(g, maxstep):
    """One instrumented run (hipEvents around every hot-path launch, which
    runs the step-by-step loop): {kernel: (launches, total ms)}."""
    g.kernel_timing(True)
    g.run(maxstep=maxstep, tol=1e-30)
    ks = g.kernel_stats()
    g.kernel_timing(False)
    return {k: v for k, v in ks.items() if v[0] > 0}


def per_kernel_table(ks, ab):
    return {k: {"launches": v[0], "avg_ms": v[1] / v[0],
                "GB/s": (ab[k] / (v[1] / v[0] / 1000.0) / 1e9) if k in ab else None,
                "algorithmic_GB": ab[k] / 1e9 if k in ab else None}
            for k, v in ks.items() if v[0]}


def phase_headline(args, budget, R):
    rism, grid, box, solv, sol, xvk, desc = workload(args.workload)
    natv = solv["natv"]
    N = grid[0] * grid[1] * grid[2]
    am = args.anderson_m
    clos = closure_code(rism, args, "PSE3" if args.workload == "cfg3" else "KH")
    log(f"workload {args.workload}: grid {grid}, natu {len(sol['sig'])}, natv {natv}")
    g = rism.RISM3D(grid, box, solv, solu_dict(sol), clos, 2, args.beta, 0.2, anderson_m=am)
    nw, nt, wall, res = timed_solves(g, args.warmup, args.steps, args.warm_cap, args.timed_cap,
                                     args.maxstep, "fp64")
    th, tm = res[-1]
    if tm["status"] != rism.STATUS_CONVERGED:
        log(f"WARNING: solve did not converge (status {tm['status']})")
    s_per = wall / nt
    iters = tm["steps"]
    mh = am if am > 0 else 2
    iter_ms = tm["iterate_ms"] / max(iters, 1)
    R["out"].update({
        "value": s_per, "steps": nt, "warmup": nw, "ms_per_step": 1000.0 * s_per,
    })
    R["out"]["config"].update({
        "workload": desc, "grid": list(grid), "box": list(box), "natu": int(len(sol["sig"])),
        "natv": int(natv), "closure": args.closure or ("PSE3" if args.workload == "cfg3"
                                                        else "KH"),
        "anderson_m": am, "iterations": iters, "status": tm["status"],
        "setup_ms": tm["setup_ms"], "iterate_ms": tm["iterate_ms"],
        "thermo_ms": tm["thermo_ms"], "ms_per_iteration": iter_ms,
        "iteration_fused_bound_GB": fused_bytes(natv, N, mh, 8) / 1e9,
        "iteration_fused_bound_frac": fused_bytes(natv, N, mh, 8) / (iter_ms / 1000.0)
        / (HBM_PEAK_GBS * 1e9),
        "SFE_SC": th["SFE_SC"], "SFE_GF": th["SFE_GF"], "PMV": th["PMV"],
        "solves_s": [r[1]["setup_ms"] / 1e3 + r[1]["iterate_ms"] / 1e3 + r[1]["thermo_ms"] / 1e3
                     for r in res],
        "steps_requested": args.steps, "warmup_requested": args.warmup,
    })
    log(f"fp64: {s_per:.3f} s/solve over {nt} solves ({iters} iterations, "
        f"{iter_ms:.2f} ms/iteration)")
    R["out"]["config"]["autotune"] = tune_summary(rism, grid)
    ctx = {"rism": rism, "grid": grid, "box": box, "solv": solv, "sol": sol, "xvk": xvk,
           "natv": natv, "N": N, "clos": clos, "am": am, "th": th, "tm": tm, "iters": iters}
    # h of the FP64 solve for rms(dh) of the mixed one, and one instrumented
    # run for the per-kernel averages
    if args.precision in ("mixed", "both"):
        ctx["h64"] = g.fields(rism.FIELD_H)
    ks = kernel_profile(g, args.profile_steps)
    g.close()
    ab = algorithmic_bytes(natv, N, mh)
    R["out"]["config"]["kernels"] = per_kernel_table(ks, ab)
    hot = {k: v for k, v in ks.items() if k in ab}
    dom = max(hot, key=lambda k: hot[k][1])
    avg_s = hot[dom][1] / hot[dom][0] / 1000.0
    achieved = ab[dom] / avg_s / 1e9
    R["out"]["roofline"] = {
        "kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
        "algorithmic_bytes_per_launch": ab[dom], "avg_launch_ms": avg_s * 1000.0,
        "source": ("hipEvents on the solver's stream around every launch of an instrumented "
                   f"run of the same input ({hot[dom][0]} launches)"),
    }
    # setup kernels: pair evaluations per second
    natu = len(sol["sig"])
    if "potential_r" in ks:
        pot_ms = ks["potential_r"][1] / ks["potential_r"][0]
        R["out"]["config"]["potential_pairs_per_s"] = N * natu / (pot_ms / 1e3)
    ctx["dom"] = dom
    return ctx


def tune_summary(rism, grid=None):
    """The autotuner's plan per tuned pass (of this grid, or all grids):
    mean launch ms of each plan and the gain of the chosen plan over the
    built-in default (plan 0)."""
    out = []
    for e in rism.tune_report():
        if grid is not None and e["grid"] != list(grid):
            continue
        ms = e["ms"]
        out.append({"pass": e["pass"], "grid": e["grid"], "dtype": e["dtype"], "plan": e["choice"],
                    "plan_ms": ms, "gain_vs_default": ms[0] / ms[e["choice"]]})
    return out


def phase_mixed(args, budget, R, ctx):
    rism = ctx["rism"]
    gm = rism.RISM3D(ctx["grid"], ctx["box"], ctx["solv"], solu_dict(ctx["sol"]), ctx["clos"],
                     2, args.beta, 0.2, anderson_m=ctx["am"], precision=1)
    nw, nt, wall, res = timed_solves(gm, 1, args.steps, 0.0, args.mixed_cap, args.maxstep,
                                     "mixed")
    thm, tmm = res[-1]
    rms_dh = None
    if ctx.get("h64") is not None:
        h32 = gm.fields(rism.FIELD_H)
        h32 -= ctx["h64"]
        rms_dh = float(np.sqrt(np.mean(h32 * h32)))
        del h32
    ksm = kernel_profile(gm, args.profile_steps)
    gm.close()
    th = ctx["th"]
    dmu = thm["SFE_SC"] - th["SFE_SC"]
    s64 = R["out"]["value"]
    N, natv = ctx["N"], ctx["natv"]
    mh = ctx["am"] if ctx["am"] > 0 else 2
    ab32 = algorithmic_bytes(natv, N, mh, w=4)
    R["out"]["mixed"] = {
        "s_per_solve": wall / nt, "solves_timed": nt, "warmup": nw,
        "speedup_vs_fp64": s64 / (wall / nt),
        # per iteration (iterate phase over steps): independent of how many
        # steps each solve's trajectory takes
        "iteration_speedup_vs_fp64": ((ctx["tm"]["iterate_ms"] / ctx["tm"]["steps"])
                                      / (tmm["iterate_ms"] / tmm["steps"])),
        "iterations": tmm["steps"], "iterations_fp32": tmm["steps_fp32"],
        "status": tmm["status"], "setup_ms": tmm["setup_ms"], "iterate_ms": tmm["iterate_ms"],
        "SFE_SC": thm["SFE_SC"], "dSFE_SC_J_per_mol": dmu, "dSFE_SC_kcal_per_mol": dmu / KCAL,
        "gate_0.05_kcal_per_mol": bool(abs(dmu) <= 0.05 * KCAL),
        "dPMV": thm["PMV"] - th["PMV"], "rms_dh": rms_dh,
        "iteration_fused_bound_frac": ((tmm["steps_fp32"] * fused_bytes(natv, N, mh, 4)
                                        + (tmm["steps"] - tmm["steps_fp32"])
                                        * fused_bytes(natv, N, mh, 8))
                                       / (tmm["iterate_ms"] / 1000.0) / (HBM_PEAK_GBS * 1e9)),
        "fp32_kernels": per_kernel_table(ksm, ab32),
    }
    log(f"mixed: {wall / nt:.3f} s/solve, {s64 / (wall / nt):.2f}x, dSFE {dmu:.3f} J/mol")


def phase_anderson(args, budget, R, ctx, m=10):
    rism = ctx["rism"]
    out = {}
    for prec, label in ((0, "fp64"), (1, "mixed")):
        if budget.left() < 60:
            out[label] = "skipped: budget"
            break
        g = rism.RISM3D(ctx["grid"], ctx["box"], ctx["solv"], solu_dict(ctx["sol"]), ctx["clos"],
                        2, args.beta, 0.2, anderson_m=m, precision=prec)
        sync()
        t0 = time.perf_counter()
        th, tm = g.run(maxstep=args.maxstep, tol=1e-6)
        sync()
        wall = time.perf_counter() - t0
        g.close()
        out[label] = {"s_per_solve": wall, "iterations": tm["steps"], "status": tm["status"],
                      "iterations_fp32": tm["steps_fp32"], "SFE_SC": th["SFE_SC"],
                      "ms_per_iteration": tm["iterate_ms"] / max(1, tm["steps"])}
        log(f"anderson({m}) {label}: {wall:.2f} s, {tm['steps']} steps")
    if isinstance(out.get("fp64"), dict) and isinstance(out.get("mixed"), dict):
        d = out["mixed"]["SFE_SC"] - out["fp64"]["SFE_SC"]
        out["mixed_speedup_vs_fp64"] = out["fp64"]["s_per_solve"] / out["mixed"]["s_per_solve"]
        out["dSFE_SC_kcal_per_mol"] = d / KCAL
    R["out"]["anderson%d" % m] = out


def phase_test2(args, budget, R):
    """The reference's own production input, timed beside the synthetic one."""
    rism, grid, box, solv, sol, xvk, desc = workload("test2")
    out = {"workload": desc}
    for prec, label in ((0, "fp64"), (1, "mixed")):
        if budget.left() < 45:
            out[label] = "skipped: budget"
            break
        g = rism.RISM3D(grid, box, solv, solu_dict(sol), rism.KH, 2, 0.6, 0.2, precision=prec)
        sync()
        t0 = time.perf_counter()
        th, tm = g.run(maxstep=args.maxstep, tol=1e-6)
        sync()
        wall = time.perf_counter() - t0
        g.close()
        out[label] = {"s_per_solve": wall, "iterations": tm["steps"], "status": tm["status"],
                      "iterations_fp32": tm["steps_fp32"], "setup_ms": tm["setup_ms"],
                      "SFE_SC": th["SFE_SC"], "SFE_GF": th["SFE_GF"], "PMV": th["PMV"]}
        log(f"test2 {label}: {wall:.2f} s, {tm['steps']} steps, SFE_SC {th['SFE_SC']:.3f}")
    if isinstance(out.get("fp64"), dict):
        out["reference_published_s"] = REF_PUBLISHED_TEST2_S
        out["reference_published_hw"] = "4 GPUs of one Cygnus node, GPUDirect (README.md:25)"
        out["speedup_vs_reference_published"] = REF_PUBLISHED_TEST2_S / out["fp64"]["s_per_solve"]
        out["reference_xmu_older_build"] = REF_TEST2_XMU
    if isinstance(out.get("mixed"), dict) and isinstance(out.get("fp64"), dict):
        d = out["mixed"]["SFE_SC"] - out["fp64"]["SFE_SC"]
        out["mixed_speedup_vs_fp64"] = out["fp64"]["s_per_solve"] / out["mixed"]["s_per_solve"]
        out["mixed_dSFE_SC_kcal_per_mol"] = d / KCAL
    R["out"]["protein_test2"] = out


def batch_cases(rism, n):
    import gen_inputs as G
    cases = []
    for i in range(n):
        sol = G.batch_solute(i)
        grid, box = G.batch_grid(sol["xyz"])
        cases.append((grid, box, solu_dict(centred(sol))))
    return cases


def phase_batch(args, budget, R, nsol=256, prec=0, stepwise=True):
    import __graft_entry__ as ge
    rism = ge.load_rism3d()
    solv = rism.read_xvk(os.path.join(GOLDEN, "tip3p_drism.xvk"))
    cases = batch_cases(rism, nsol)

    def make():
        return [rism.RISM3D(g, b, solv, su, rism.KH, 2, 0.6, 0.2, precision=prec)
                for g, b, su in cases]

    w = make()[:4]  # warm-up (first launches of the small-grid kernels)
    rism.run_batch(w, threads=args.batch_threads)
    for g in w:
        g.close()
    gs = make()
    sync()
    t0 = time.perf_counter()
    out = rism.run_batch(gs, threads=args.batch_threads)
    sync()
    t_batch = time.perf_counter() - t0
    for g in gs:
        g.close()
    steps = np.array([o[2] for o in out])
    sfe_batch = [o[0]["SFE_SC"] for o in out]
    gs = make()
    sync()
    t0 = time.perf_counter()
    sfe_seq = [g.run()[0]["SFE_SC"] for g in gs]
    sync()
    t_seq = time.perf_counter() - t0
    for g in gs:
        g.close()
    res = {
        "workload": ("%d random-walk solutes (8-60 atoms, seeds 1004000+i, net charge -1/0/+1) "
                     "in TIP3P water (chi from the project's 1D DRISM-KH, eps 78.4), boxes = "
                     "extent + 12 A/side at 0.5 A, grids from {48..96}^3, KH, AN2 2 0.6 0.2, "
                     "tol 1e-6, %s" % (nsol, "mixed precision" if prec else "FP64")),
        "solves": nsol, "grid_shapes": len(set(c[0] for c in cases)),
        "solves_per_s": nsol / t_batch, "batch_s": t_batch,
        "sequential_s": t_seq, "sequential_solves_per_s": nsol / t_seq,
        "speedup_vs_sequential": t_seq / t_batch,
        "iterations_p50": float(np.percentile(steps, 50)),
        "iterations_p99": float(np.percentile(steps, 99)),
        "converged": int(sum(1 for o in out if o[1] == rism.STATUS_CONVERGED)),
        "batch_equals_sequential": bool(all(a == b for a, b in zip(sfe_batch, sfe_seq))),
        "host_threads": args.batch_threads,
    }
    if stepwise and budget.left() > 60:
        # the same solves one after another with a host round trip per step
        # (the reference's loop structure: rms read back before every update)
        gs = make()
        old = os.environ.get("RISM_SYNC_FREE")
        os.environ["RISM_SYNC_FREE"] = "0"  # rism_solve's step-by-step loop
        try:
            sync()
            t0 = time.perf_counter()
            for g in gs:
                g.run()
            sync()
            t_step = time.perf_counter() - t0
        finally:
            if old is None:
                del os.environ["RISM_SYNC_FREE"]
            else:
                os.environ["RISM_SYNC_FREE"] = old
        for g in gs:
            g.close()
        res["sequential_stepwise_s"] = t_step
        res["speedup_vs_sequential_stepwise"] = t_step / t_batch
    res["autotune"] = tune_digest(rism, set(tuple(c[0]) for c in cases))
    log(f"batch: {nsol / t_batch:.1f} solves/s, {t_seq / t_batch:.2f}x sequential")
    return res


def tune_digest(rism, grids):
    """The autotuner's choices over a set of grids, per tuned pass and bulk
    type: how many grids it timed, how many it moved off the default plan,
    and the mean and largest gain of the chosen plan over the default."""
    agg = {}
    for e in rism.tune_report():
        if tuple(e["grid"]) not in grids:
            continue
        ms = e["ms"]
        gain = ms[0] / ms[e["choice"]]
        a = agg.setdefault((e["pass"], e["dtype"]), [0, 0, 0.0, 1.0])
        a[0] += 1
        a[1] += 1 if e["choice"] != 0 else 0
        a[2] += gain
        a[3] = max(a[3], gain)
    return [{"pass": k[0], "dtype": k[1], "grids": v[0], "off_default": v[1],
             "gain_mean": v[2] / v[0], "gain_max": v[3]} for k, v in sorted(agg.items())]


# ---------------------------------------------------------------------------
# HBM traffic from rocprofv3 counters (child processes)
# ---------------------------------------------------------------------------
COPY_BYTES = 2 << 30


def kernel_key(name):
    """Maps a dispatched kernel's name to the solver's kernel names."""
    if "k_copy_stream" in name:
        return "copy"
    if "k_zfwd_closure" in name:
        return "zfwd_closure"
    if "k_xpass_oz" in name:
        return "xpass_oz"
    if "k_zinv_resid" in name:
        return "zinv_resid"
    if "k_an2_update" in name:
        return "an2_update"
    if "k_andm_update" in name:
        return "an2_update"
    if "k_ypass<" in name:
        args = name.split("k_ypass<", 1)[1].split(">", 1)[0].split(",")
        return "ypass_fwd" if args[2].strip() == "-1" else "ypass_inv"
    return None


def read_counters(outdir, counter):
    """{kernel key: [value per dispatch]} from rocprofv3 csv output, for each
    key the instantiation with the most dispatches: the solver step's kernel
    in the chosen launch plan (the autotuner's trials of the other plans and
    the Picard-step variants run a few times each)."""
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    per = {}
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                name = row.get("Kernel_Name", "")
                k = kernel_key(name)
                if k:
                    per.setdefault(k, {}).setdefault(name, []).append(float(row["Counter_Value"]))
    return {k: max(d.values(), key=len) for k, d in per.items()}


def phase_traffic(args, budget, R, ctx):
    """HBM bytes per launch of the hot-loop kernels from two rocprofv3 --pmc
    child runs (FETCH_SIZE, then WRITE_SIZE), each scaled by what the counter
    reads for a plain copy of COPY_BYTES in the same run."""
    rp = shutil.which("rocprofv3")
    if not rp:
        return "rocprofv3 not found"
    base = os.path.join(ROOT, "gpurun_out", "bench_traffic")
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        if budget.left() < 40:
            return f"skipped: budget ({counter})"
        outdir = os.path.join(base, counter.lower())
        shutil.rmtree(outdir, ignore_errors=True)
        cmd = ["timeout", "-s", "KILL", "120", rp, "--pmc", counter, "--output-format", "csv",
               "-d", outdir, "-o", "trace", "--", sys.executable, os.path.abspath(__file__),
               "--profile-run", args.workload, "--profile-steps", "8",
               "--anderson-m", str(args.anderson_m)]
        log("traffic:", " ".join(cmd[4:]))
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        vals = read_counters(outdir, counter)
        if "copy" not in vals or not vals["copy"]:
            return f"{counter}: no counter data (rc {p.returncode}): {p.stdout[-400:]}"
        res[counter] = vals
    fcal = COPY_BYTES / (np.mean(res["FETCH_SIZE"]["copy"]) * 1024.0)
    wcal = COPY_BYTES / (np.mean(res["WRITE_SIZE"]["copy"]) * 1024.0)
    table = {}
    for k in res["FETCH_SIZE"]:
        if k == "copy" or k not in res["WRITE_SIZE"]:
            continue
        f, w = res["FETCH_SIZE"][k], res["WRITE_SIZE"][k]
        # the last launches: the solver's steps (the first are the autotuner's)
        f, w = f[len(f) // 2:], w[len(w) // 2:]
        rd = float(np.mean(f)) * 1024.0 * fcal
        wr = float(np.mean(w)) * 1024.0 * wcal
        table[k] = {"read_GB": rd / 1e9, "write_GB": wr / 1e9, "traffic_GB": (rd + wr) / 1e9}
    return {"kernels": table, "fetch_factor": fcal, "write_factor": wcal,
            "calibration": f"k_copy_stream of {COPY_BYTES} bytes in the same run"}


def profile_run(args):
    """Child of phase_traffic (and of the committed rocprofv3 summaries): a
    plain copy of COPY_BYTES three times, then the set-up and a few
    instrumented (step-by-step) steps of the workload."""
    rism, grid, box, solv, sol, xvk, desc = workload(args.profile_run)
    ms = rism.copy_bench(COPY_BYTES, 3)
    print(f"copy {COPY_BYTES} B: {ms:.3f} ms ({2 * COPY_BYTES / ms / 1e6:.0f} GB/s)", flush=True)
    prec = 1 if args.precision == "mixed" else 0
    clos = closure_code(rism, args, "PSE3" if args.profile_run == "cfg3" else "KH")
    g = rism.RISM3D(grid, box, solv, solu_dict(sol), clos, 2, args.beta, 0.2,
                    anderson_m=args.anderson_m, precision=prec)
    g.kernel_timing(True)
    th, tm = g.run(maxstep=args.profile_steps, tol=1e-30)
    ks = g.kernel_stats()
    g.close()
    print(json.dumps({"profile_run": args.profile_run, "steps": tm["steps"],
                      "kernels_avg_ms": {k: v[1] / v[0] for k, v in ks.items() if v[0]}}),
          flush=True)


# ---------------------------------------------------------------------------
# CPU baseline (the oracle on the host cores)
# ---------------------------------------------------------------------------
def cpu_baseline(grid, box, solv_path, sol, gpu_steps, natv):
    """FP64 oracle on the host cores, on a bounded sample extrapolated to one
    solve: one oracle step + one AN2 update at the full grid (synthetic
    fields: the cost does not depend on the values), plus LJ/Coulomb and fk
    on a slab of rows scaled to the full grid."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    L = O.lib()
    threads = L.or_num_threads()
    s = O.read_xvk(solv_path)
    nx, ny, nz = grid
    N = nx * ny * nz
    rows = 16
    sub = (nx, rows, 1)
    xyz = np.ascontiguousarray(sol["xyz"], dtype=np.float64)
    so = O.Solute(np.asarray(sol["sig"], float), np.asarray(sol["eps"], float),
                  np.asarray(sol["q"], float), xyz)
    t0 = time.time()
    O.potentials(sub, (box[0], box[1] * rows / ny, box[2] / nz), s, so)
    t_pot = (time.time() - t0) * N / (rows * nx)
    ga = np.linspace(0.01, 10.0, 64)
    chi = np.tile(np.eye(natv).reshape(-1), len(ga))
    indga = np.zeros(N, dtype=np.int32)
    u = np.zeros(natv * N)
    fr = np.zeros(N)
    fk = np.zeros(2 * N)
    t = np.zeros(natv * N)
    g = np.zeros(natv * N)
    R = np.zeros(natv * N)
    h = np.zeros(natv * N)
    work = np.zeros(4 * natv * N)
    t0 = time.time()
    L.or_calculate(nx, ny, nz, *[float(b) for b in box], natv, 0, np.ascontiguousarray(s.q), t,
                   u, fr, fk, chi, indga, g, R, h, work)
    t_step = time.time() - t0
    del work, g, h
    A = L.or_an2_create(natv, N, np.ascontiguousarray(s.rho), 0, 0.6, 0.2)
    t0 = time.time()
    L.or_an2_step(A, t, R)
    t_an2 = time.time() - t0
    L.or_an2_destroy(A)
    per_solve = t_pot + gpu_steps * (t_step + t_an2)
    return {
        "value": per_solve, "unit": "s/solve", "cores": int(threads), "kind": "port",
        "sample": (f"FP64 oracle (oracle/rism_oracle.c, OpenMP, own mixed-radix FFT): LJ+Coulomb"
                   f"+fk on {rows} rows x {nx} of one plane scaled x{N // (rows * nx)} = "
                   f"{t_pot:.1f} s; one step + one AN2 update at {nx}x{ny}x{nz} = "
                   f"{t_step + t_an2:.1f} s, x {gpu_steps} steps (the GPU's count); extrapolated"),
    }


def cpu_full_cfg1():
    """A full FP64 oracle solve of the 64^3 case, beside the GPU's time for the
    same solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import gen_inputs as G
    sol = centred(G.cfg1_solute())
    s = O.read_xvk(os.path.join(GOLDEN, "tip3p_drism.xvk"))
    so = O.Solute(np.asarray(sol["sig"], float), np.asarray(sol["eps"], float),
                  np.asarray(sol["q"], float), np.ascontiguousarray(sol["xyz"], float))
    t0 = time.time()
    o = O.OracleRISM((64, 64, 64), (32.0, 32.0, 32.0), s, so, "KH", 2, 0.6, 0.2, 1e-6,
                     10000).setup()
    r = o.solve()
    th = o.thermo()
    return {"s_per_solve": time.time() - t0, "iterations": r.steps, "SFE_SC": th["SFE_SC"]}


def gpu_cfg1(args):
    rism, grid, box, solv, sol, xvk, desc = workload("cfg1")
    g = rism.RISM3D(grid, box, solv, solu_dict(sol), rism.KH, 2, 0.6, 0.2)
    g.run()
    sync()
    t0 = time.perf_counter()
    n = 5
    for _ in range(n):
        th, tm = g.run()
    sync()
    wall = (time.perf_counter() - t0) / n
    g.close()
    return {"s_per_solve": wall, "iterations": tm["steps"], "SFE_SC": th["SFE_SC"]}


def closure_code(rism, args, default):
    name = args.closure or default
    return {"KH": rism.KH, "HNC": rism.HNC, "PSE3": rism.PSE(3)}[name]


# ---------------------------------------------------------------------------
# other workloads (not the headline line)
# ---------------------------------------------------------------------------
def bench_medium(args):
    """Batched medium-size solutes in mixed precision: 32 protein-like
    clusters (500-4000 atoms log-uniform, 0.1 atoms/A^3, net charge 0, seeds
    1005000 + i) in 0.2 M NaCl TIP3P (tip3p_nacl02_drism.xvk), boxes of
    extent + 12 A per side at 0.5 A spacing with per-axis sizes from
    {128, 144, 160, 180, 192}; the configured HNC closure by default."""
    import gen_inputs as G
    import __graft_entry__ as ge
    rism = ge.load_rism3d()
    solv = rism.read_xvk(os.path.join(GOLDEN, "tip3p_nacl02_drism.xvk"))
    nsol = args.batch_size if args.batch_size else 32
    clos = closure_code(rism, args, "HNC")
    cname = args.closure or "HNC"
    cases, natoms = [], []
    for i in range(nsol):
        sol = G.medium_solute(i)
        grid, box = G.batch_grid(sol["xyz"], sizes=G.MEDIUM_SIZES)
        cases.append((grid, box, solu_dict(centred(sol))))
        natoms.append(len(sol["sig"]))
    log(f"batch-medium: {nsol} solutes ({min(natoms)}-{max(natoms)} atoms), "
        f"{len(set(c[0] for c in cases))} grid shapes")

    def make(prec):
        return [rism.RISM3D(g, b, solv, su, clos, 2, 0.6, 0.2, precision=prec)
                for g, b, su in cases]

    def timed_batch(prec):
        gs = make(prec)
        sync()
        t0 = time.perf_counter()
        out = rism.run_batch(gs, threads=args.batch_threads)
        sync()
        t = time.perf_counter() - t0
        for g in gs:
            g.close()
        return t, out

    w = make(1)[:2]
    rism.run_batch(w, threads=args.batch_threads)
    for g in w:
        g.close()
    t_mixed, out_m = timed_batch(1)
    gs = make(1)
    sync()
    t0 = time.perf_counter()
    seq = [g.run()[0]["SFE_SC"] for g in gs]
    sync()
    t_seq = time.perf_counter() - t0
    for g in gs:
        g.close()
    t_fp64, out_d = timed_batch(0)
    conv = [i for i in range(nsol) if out_m[i][1] == rism.STATUS_CONVERGED
            and out_d[i][1] == rism.STATUS_CONVERGED]
    dmu = np.array([abs(out_m[i][0]["SFE_SC"] - out_d[i][0]["SFE_SC"]) / KCAL for i in conv])
    steps = np.array([o[2] for o in out_m])
    res = base_line(args)
    res.update({
        "value": nsol / t_mixed, "unit": "solves/s", "steps": 1, "warmup": 1,
        "ms_per_step": 1000.0 * t_mixed, "higher_is_better": True,
        "dtype": "f32 bulk, f64 sums and FP64 steps",
    })
    res["config"] = {
        "workload": ("batch-medium: %d protein-like clusters (500-4000 atoms, seeds 1005000+i, "
                     "0 e) in 0.2 M NaCl TIP3P (chi from the project's 1D DRISM-KH, eps 78.4), "
                     "boxes = extent + 12 A/side at 0.5 A, grids from {128..192}^3, %s, AN2 "
                     "2 0.6 0.2, tol 1e-6, mixed precision" % (nsol, cname)),
        "solves": nsol, "grid_shapes": len(set(c[0] for c in cases)),
        "atoms_min_max": [min(natoms), max(natoms)],
        "batch_s": t_mixed, "sequential_s": t_seq, "speedup_vs_sequential": t_seq / t_mixed,
        "batch_equals_sequential": bool(all(out_m[i][0]["SFE_SC"] == seq[i]
                                            for i in range(nsol))),
        "fp64_batch_s": t_fp64, "speedup_vs_fp64_batch": t_fp64 / t_mixed,
        "iterations_p50": float(np.percentile(steps, 50)),
        "iterations_p99": float(np.percentile(steps, 99)),
        "converged_mixed": int(sum(1 for o in out_m if o[1] == rism.STATUS_CONVERGED)),
        "converged_fp64": int(sum(1 for o in out_d if o[1] == rism.STATUS_CONVERGED)),
        "dSFE_kcal_per_mol_max": float(dmu.max()) if len(dmu) else None,
        "dSFE_kcal_per_mol_p50": float(np.percentile(dmu, 50)) if len(dmu) else None,
        "within_0.05_kcal_per_mol": int((dmu <= 0.05).sum()),
        "autotune": tune_digest(rism, set(tuple(c[0]) for c in cases)),
    }
    print(json.dumps(res), flush=True)


def base_line(args):
    return {
        "metric": METRIC, "value": None, "unit": "s/solve", "n_gpus": 1, "steps": 0,
        "warmup": 0, "ms_per_step": None, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "budget_s": args.budget},
        "roofline": None, "cpu_baseline": None,
    }


# ---------------------------------------------------------------------------
# main
# ---------------------------------------------------------------------------
_print_lock = threading.Lock()
_printed = [False]


def emit(obj):
    with _print_lock:
        if _printed[0]:
            return
        _printed[0] = True
        print(json.dumps(obj), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed solves (at most; the budget may stop earlier)")
    ap.add_argument("--warmup", type=int, default=1,
                    help="untimed warm-up solves (at most; at least one runs)")
    ap.add_argument("--budget", type=float, default=450.0,
                    help="wall-clock seconds from process start for the whole run")
    ap.add_argument("--workload", default="cfg2",
                    choices=["cfg2", "cfg1", "cfg3", "mobley", "test2", "batch", "batch-medium"])
    ap.add_argument("--closure", default=None, choices=["KH", "HNC", "PSE3"],
                    help="override the workload's closure (cfg3: PSE3, batch-medium: HNC, "
                         "others: KH)")
    ap.add_argument("--batch-size", type=int, default=None)
    ap.add_argument("--batch-threads", type=int, default=8)
    ap.add_argument("--anderson-m", type=int, default=0,
                    help="history length of Anderson(m) mixing for the headline (0: AN2)")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--beta", type=float, default=0.6,
                    help="mixing coefficient (AN2's m; beta of Anderson(m))")
    ap.add_argument("--only", default=None,
                    help="comma list of extra phases to run (mixed,batch,test2,traffic,cpu,"
                         "anderson10); default all, as the budget allows")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--maxstep", type=int, default=10000)
    ap.add_argument("--profile-steps", type=int, default=40,
                    help="steps of the instrumented per-kernel run")
    ap.add_argument("--warm-cap", type=float, default=20.0)
    ap.add_argument("--timed-cap", type=float, default=90.0)
    ap.add_argument("--mixed-cap", type=float, default=45.0)
    ap.add_argument("--profile-run", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs one solve per GPU and has no "
              "multi-GPU path; run with --gpus 1", flush=True)
        sys.exit(2)
    if device_count() < 1:
        raise SystemExit("bench.py needs a GPU")
    if hip().hipSetDevice(0) != 0:
        raise SystemExit("hipSetDevice(0) failed")
    if args.profile_run:
        profile_run(args)
        return
    if args.workload == "batch-medium":
        bench_medium(args)
        return

    budget = Budget(args.budget)
    R = {"out": base_line(args)}
    # a last-resort line if the run overstays its budget (a slow box): the
    # phases measured so far
    hard = max(30.0, min(args.budget + 90.0, 570.0) - (time.time() - T_START))

    def watchdog():
        R["out"]["truncated"] = "budget exceeded; later phases not run"
        emit(R["out"])
        os._exit(0)

    wd = threading.Timer(hard, watchdog)
    wd.daemon = True
    wd.start()

    if args.workload == "batch":
        res = phase_batch(args, budget, R, nsol=args.batch_size or 256,
                          prec=1 if args.precision == "mixed" else 0)
        R["out"].update({"value": res["solves_per_s"], "unit": "solves/s", "steps": 1,
                         "warmup": 1, "ms_per_step": 1000.0 * res["batch_s"],
                         "higher_is_better": True})
        R["out"]["config"].update(res)
        emit(R["out"])
        return

    want = set((args.only or "mixed,batch,test2,traffic,cpu,anderson10").split(","))
    ctx = phase_headline(args, budget, R)
    log(f"headline done, {budget.left():.0f} s of budget left")
    phases = []
    if args.precision in ("mixed", "both") and "mixed" in want:
        phases.append(("mixed", 30 + 1.6 * R["out"]["value"],
                       lambda: phase_mixed(args, budget, R, ctx)))
    if args.workload == "cfg2" and "batch" in want:
        phases.append(("batch", 40, lambda: R["out"].__setitem__(
            "batch", phase_batch(args, budget, R))))
    if args.workload == "cfg2" and "test2" in want:
        phases.append(("test2", 50, lambda: phase_test2(args, budget, R)))
    if "traffic" in want:
        def run_traffic():
            tr = phase_traffic(args, budget, R, ctx)
            R["out"]["hbm_traffic"] = tr
            if isinstance(tr, dict) and ctx["dom"] in tr["kernels"]:
                R["out"]["roofline"]["traffic"] = tr["kernels"][ctx["dom"]]["traffic_GB"] * 1e9
        phases.append(("traffic", 45, run_traffic))
    if not args.no_cpu and "cpu" in want:
        def run_cpu():
            cb = cpu_baseline(ctx["grid"], ctx["box"], os.path.join(GOLDEN, ctx["xvk"]),
                              ctx["sol"], ctx["iters"], ctx["natv"])
            if budget.left() > 60:
                c1 = cpu_full_cfg1()
                g1 = gpu_cfg1(args)
                cb["full_solve_64^3"] = {"cpu": c1, "gpu": g1,
                                         "gpu_speedup": c1["s_per_solve"] / g1["s_per_solve"]}
            R["out"]["cpu_baseline"] = cb
        phases.append(("cpu", 40, run_cpu))
    if args.workload == "cfg2" and "anderson10" in want:
        phases.append(("anderson10", 60, lambda: phase_anderson(args, budget, R, ctx)))
    skipped = []
    for name, need, fn in phases:
        if budget.left() < need:
            skipped.append(name)
            log(f"skip {name}: {budget.left():.0f} s left, needs ~{need:.0f} s")
            continue
        t0 = time.time()
        try:
            fn()
        except Exception as e:  # a failed side phase must not lose the headline
            R["out"].setdefault("errors", {})[name] = repr(e)[:400]
            log(f"phase {name} failed: {e!r}")
        log(f"phase {name}: {time.time() - t0:.1f} s, {budget.left():.0f} s left")
    if skipped:
        R["out"]["skipped_phases"] = skipped
    R["out"]["config"]["wall_s"] = time.time() - T_START
    wd.cancel()
    emit(R["out"])


if __name__ == "__main__":
    main()
