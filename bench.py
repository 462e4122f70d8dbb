"""Benchmark: seconds per converged 3D-RISM solve on one MI355X.

Workload (default "cfg2"): 512^3 grid in a 256 A box; a synthetic 12877-atom
protein-like cluster (uniform in a 30 A sphere, >= 1.0 A apart, C/N/O/H + 1 %
S, charges ~ N(0, 0.35 e) neutralised to +2 e, seed 1002); 0.2 M NaCl in
TIP3P water (chi_vv from the project's 1D DRISM-KH at eps 78.4, 298.15 K,
tests/golden/tip3p_nacl02_drism.xvk, 4 sites); KH closure; AN2 2 0.6 0.2;
tolerance 1e-6; FP64 throughout.  One step = one full solve:
LJ/Coulomb potentials, the initial t, the fixed-point loop to convergence and
the thermodynamics, with the solute and the chi table already resident in
HBM when the clock starts.

Also reported, per the roofline contract: the dominant hot-loop kernel's
achieved HBM bandwidth (its algorithmic bytes per launch / its average
hipEvent-timed launch duration, measured on the solver's own stream during
the timed solves), and an FP64 CPU baseline (the oracle, a plain-C OpenMP
restatement of the reference algorithm) timed on a bounded sample on the
host cores and extrapolated to one solve.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
                  [--workload cfg2|cfg1|cfg3|mobley|batch|batch-medium]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "3d-rism-cuda-mpi_amd"))

METRIC = ("s/solve at 512^3, batched solves/s; mixed vs FP64 at "
          "|Δμ|<0.05 kcal/mol")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    import gen_inputs as G
    import __graft_entry__ as ge
    rism = ge.load_rism3d()
    if name == "cfg2":
        sol = G.cfg2_solute()
        xvk = "tip3p_nacl02_drism.xvk"
        grid, box = (512, 512, 512), (256.0, 256.0, 256.0)
        desc = ("cfg2: 512^3 grid / 256 A box, synthetic 12877-atom cluster (R 30 A, +2 e, "
                "seed 1002), 0.2 M NaCl in TIP3P (chi from the project's 1D DRISM-KH, "
                "eps 78.4, 298.15 K, 4 sites), KH, AN2 2 0.6 0.2, tol 1e-6, FP64")
    elif name == "cfg1":
        sol = G.cfg1_solute()
        xvk = "tip3p_drism.xvk"
        grid, box = (64, 64, 64), (32.0, 32.0, 32.0)
        desc = ("cfg1: 64^3 / 32 A, 24-atom random-walk solute (seed 1001), TIP3P water "
                "(chi from the project's 1D DRISM-KH, eps 78.4, 298.15 K), KH, AN2 2 0.6 0.2, "
                "tol 1e-6, FP64")
    elif name == "cfg3":
        sol = G.cfg3_solute()
        xvk = "tip3p_kcl01_drism.xvk"
        grid, box = (256, 256, 256), (128.0, 128.0, 128.0)
        desc = ("cfg3: 256^3 / 128 A, synthetic 3000-atom cluster (R 19 A, 0 e, seed 1003), "
                "0.1 M KCl in TIP3P (chi from the project's 1D DRISM-KH, eps 78.4, 298.15 K, "
                "4 sites), tol 1e-6, FP64")
    elif name == "mobley":
        d = rism.read_input(os.path.join(GOLDEN, "3drism.dat"),
                            os.path.join(GOLDEN, "mobley_3053621"), solvent_dir=GOLDEN)
        sol = {"sig": d["usig"], "eps": d["ueps"], "q": d["uq"], "xyz": d["uxyz"]}
        xvk = "ctip3p.xvk"
        grid, box = (256, 256, 256), (64.0, 64.0, 64.0)
        desc = "example_milc mobley_3053621: 256^3 / 64 A, 12 atoms, ctip3p.xvk, KH, FP64"
    else:
        raise SystemExit(f"unknown workload {name}")
    if name != "mobley":
        from_centre = sol["xyz"].copy()
        mn, mx = from_centre.min(0), from_centre.max(0)
        shift = np.array([np.floor(-(mx[i] - mn[i]) / 2 - mn[i] + 0.5) for i in range(3)])
        sol = dict(sol, xyz=from_centre + shift)  # solute.cu:18-48 centring
    solv = rism.read_xvk(os.path.join(GOLDEN, xvk))
    return rism, grid, box, solv, sol, xvk, desc


# HBM bytes per launch of the hot-loop kernels at cfg2 (512^3, 4 sites, FP64),
# measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (one counter per
# run) and scaled by the same counters on a plain 2 GiB device copy
# (fetch x2, write x1).
MEASURED_TRAFFIC_CFG2 = {
    "zfwd_closure": 13.97e9, "ypass_fwd": 8.59e9, "xpass_oz": 11.73e9, "ypass_inv": 8.59e9,
    "zinv_resid": 24.35e9, "an2_update": 25.77e9,
}


def algorithmic_bytes(natv, N, hist=2):
    """Minimal HBM bytes per launch of each hot-loop kernel (FP64 words of 8 B)
    with `hist` history residuals (2 for AN2, m for Anderson(m)):
    zfwd reads t, u (2n) and fr (1), writes the half spectrum (n);
    each y pass reads and writes the spectrum (2n);
    xpass_oz reads the spectrum (n) and the half fk (1), writes (n);
    zinv reads the spectrum, t, u (3n) and the history residuals (hist n),
    writes R (n);
    the update reads t, R and the history P_p (hist n), writes t', P_new (2n)."""
    w = 8.0 * N
    n = natv
    return {
        "zfwd_closure": (3 * n + 1) * w,
        "ypass_fwd": 2 * n * w,
        "xpass_oz": (2 * n + 1) * w,
        "ypass_inv": 2 * n * w,
        "zinv_resid": (4 + hist) * n * w,
        "an2_update": (4 + hist) * n * w,
    }


def cpu_baseline(grid, box, solv_path, sol, gpu_steps, natv, budget_s=30.0):
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
    # potentials on `rows` y-rows of one z-plane, scaled by N / (rows * nx)
    rows = 16
    sub = (nx, rows, 1)
    xyz = np.ascontiguousarray(sol["xyz"], dtype=np.float64)
    so = O.Solute(np.asarray(sol["sig"], float), np.asarray(sol["eps"], float),
                  np.asarray(sol["q"], float), xyz)
    t0 = time.time()
    O.potentials(sub, (box[0], box[1] * rows / ny, box[2] / nz), s, so)
    t_pot = (time.time() - t0) * N / (rows * nx)
    # one step + one AN2 update at the full grid
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
                   f"{t_step + t_an2:.1f} s, x {gpu_steps} steps (the GPU's count)"),
    }


def closure_code(rism, args, default):
    name = args.closure or default
    return {"KH": rism.KH, "HNC": rism.HNC, "PSE3": rism.PSE(3)}[name]


def bench_medium(args):
    """Batched medium-size solutes in mixed precision: 32 protein-like
    clusters (500-4000 atoms log-uniform, 0.1 atoms/A^3, net charge 0, seeds
    1005000 + i) in 0.2 M NaCl TIP3P (tip3p_nacl02_drism.xvk), boxes of
    extent + 12 A per side at 0.5 A spacing with per-axis sizes from
    {128, 144, 160, 180, 192}; HNC, AN2 2 0.6 0.2, tol 1e-6.  Times the
    mixed-precision batch (rism_run_batch) against the same mixed solves one
    after another, and reports |dSFE| of every solute against its FP64 solve
    (run as a sync-free FP64 batch, also timed)."""
    import torch
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
        xyz = sol["xyz"]
        mn, mx = xyz.min(0), xyz.max(0)
        shift = np.array([np.floor(-(mx[k] - mn[k]) / 2 - mn[k] + 0.5) for k in range(3)])
        cases.append((grid, box, {"usig": sol["sig"], "ueps": sol["eps"], "uq": sol["q"],
                                  "uxyz": xyz + shift}))
        natoms.append(len(sol["sig"]))
    log(f"[bench] batch-medium: {nsol} solutes ({min(natoms)}-{max(natoms)} atoms), "
        f"{len(set(c[0] for c in cases))} grid shapes")

    def make(prec):
        return [rism.RISM3D(g, b, solv, su, clos, 2, 0.6, 0.2, precision=prec)
                for g, b, su in cases]

    def timed_batch(prec):
        gs = make(prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = rism.run_batch(gs, threads=args.batch_threads)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        for g in gs:
            g.close()
        return t, out

    w = make(1)[:2]  # warm-up (module load, first launches)
    rism.run_batch(w, threads=args.batch_threads)
    for g in w:
        g.close()
    t_mixed, out_m = timed_batch(1)
    log(f"[bench] mixed batch: {t_mixed:.2f} s")
    gs = make(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq = [g.run()[0]["SFE_SC"] for g in gs]
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    for g in gs:
        g.close()
    log(f"[bench] mixed one after another: {t_seq:.2f} s")
    t_fp64, out_d = timed_batch(0)
    log(f"[bench] FP64 batch: {t_fp64:.2f} s")
    conv = [i for i in range(nsol) if out_m[i][1] == rism.STATUS_CONVERGED
            and out_d[i][1] == rism.STATUS_CONVERGED]
    dmu = np.array([abs(out_m[i][0]["SFE_SC"] - out_d[i][0]["SFE_SC"]) / 4184.0 for i in conv])
    steps = np.array([o[2] for o in out_m])
    res = {
        "metric": METRIC,
        "value": nsol / t_mixed,
        "unit": "solves/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1000.0 * t_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk, f64 sums and FP64 steps",
        "data": "synthetic",
        "config": {
            "workload": ("batch-medium: %d protein-like clusters (500-4000 atoms, seeds "
                         "1005000+i, 0 e) in 0.2 M NaCl TIP3P (chi from the project's 1D "
                         "DRISM-KH, eps 78.4), boxes = extent + 12 A/side at 0.5 A, grids "
                         "from {128..192}^3, %s, AN2 2 0.6 0.2, tol 1e-6, mixed precision"
                         % (nsol, cname)),
            "solves": nsol, "grid_shapes": len(set(c[0] for c in cases)),
            "atoms_min_max": [min(natoms), max(natoms)],
            "batch_s": t_mixed, "sequential_s": t_seq,
            "speedup_vs_sequential": t_seq / t_mixed,
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
            "host_threads": args.batch_threads,
        },
    }
    print(json.dumps(res), flush=True)


def bench_batch(args):
    """Batched small solutes: 256 random-walk solutes of 8-60 atoms (seeds
    1004000 + i, net charge -1/0/+1) in TIP3P water (tip3p_drism.xvk), each in
    its own box of extent + 12 A padding per side at 0.5 A spacing, per-axis
    grid sizes from {48, 60, 64, 72, 80, 96}; KH, AN2 2 0.6 0.2, tol 1e-6,
    FP64.  Times the batch (rism_run_batch, one stream per solute, host
    threads enqueueing) against the same solves run one after another."""
    import torch
    import gen_inputs as G
    import __graft_entry__ as ge
    rism = ge.load_rism3d()
    solv = rism.read_xvk(os.path.join(GOLDEN, "tip3p_drism.xvk"))
    nsol = args.batch_size if args.batch_size else 256
    cases = []
    for i in range(nsol):
        sol = G.batch_solute(i)
        grid, box = G.batch_grid(sol["xyz"])
        xyz = sol["xyz"]
        mn, mx = xyz.min(0), xyz.max(0)
        shift = np.array([np.floor(-(mx[k] - mn[k]) / 2 - mn[k] + 0.5) for k in range(3)])
        cases.append((grid, box, {"usig": sol["sig"], "ueps": sol["eps"], "uq": sol["q"],
                                  "uxyz": xyz + shift}))
    log(f"[bench] batch: {nsol} solutes, {len(set(c[0] for c in cases))} grid shapes")

    prec = 1 if args.precision == "mixed" else 0

    def make():
        return [rism.RISM3D(g, b, solv, su, rism.KH, 2, 0.6, 0.2, precision=prec)
                for g, b, su in cases]

    # warm-up (module load, first launches)
    w = make()[:4]
    rism.run_batch(w, threads=args.batch_threads)
    for g in w:
        g.close()
    # batched
    gs = make()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = rism.run_batch(gs, threads=args.batch_threads)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    steps = np.array([o[2] for o in out])
    status = [o[1] for o in out]
    sfe_batch = [o[0]["SFE_SC"] for o in out]
    for g in gs:
        g.close()
    # the same solves one after another
    gs = make()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sfe_seq = []
    for g in gs:
        th, tm = g.run()
        sfe_seq.append(th["SFE_SC"])
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    for g in gs:
        g.close()
    same = all(a == b for a, b in zip(sfe_batch, sfe_seq))
    # the same solves one after another with a host round trip per step (the
    # reference's loop structure: rms read back before every update)
    gs = make()
    for g in gs:
        g.kernel_timing(True)  # per-kernel events: forces the step-by-step loop
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for g in gs:
        g.run()
    torch.cuda.synchronize()
    t_step = time.perf_counter() - t0
    for g in gs:
        g.close()
    res = {
        "metric": METRIC,
        "value": nsol / t_batch,
        "unit": "solves/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1000.0 * t_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk, f64 sums and FP64 steps" if prec else "f64",
        "data": "synthetic",
        "config": {
            "workload": ("batch: %d random-walk solutes (8-60 atoms, seeds 1004000+i, net "
                         "charge -1/0/+1) in TIP3P water (chi from the project's 1D DRISM-KH, "
                         "eps 78.4), boxes = extent + 12 A/side at 0.5 A, grids from "
                         "{48..96}^3, KH, AN2 2 0.6 0.2, tol 1e-6, %s"
                         % (nsol, "mixed precision" if prec else "FP64")),
            "solves": nsol, "grid_shapes": len(set(c[0] for c in cases)),
            "batch_s": t_batch, "sequential_s": t_seq,
            "sequential_solves_per_s": nsol / t_seq,
            "speedup_vs_sequential": t_seq / t_batch,
            "sequential_stepwise_s": t_step,
            "speedup_vs_sequential_stepwise": t_step / t_batch,
            "iterations_p50": float(np.percentile(steps, 50)),
            "iterations_p99": float(np.percentile(steps, 99)),
            "converged": int(sum(1 for s in status if s == rism.STATUS_CONVERGED)),
            "batch_equals_sequential": bool(same),
            "host_threads": args.batch_threads,
        },
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg2",
                    choices=["cfg2", "cfg1", "cfg3", "mobley", "batch", "batch-medium"],
                    help="cfg2 (512^3, default), cfg1 (64^3), cfg3 (256^3, PSE-3), mobley, "
                         "batch (256 small solutes) or batch-medium (32 medium solutes, "
                         "mixed precision)")
    ap.add_argument("--closure", default=None, choices=["KH", "HNC", "PSE3"],
                    help="override the workload's closure (cfg3: PSE3, batch-medium: HNC, "
                         "others: KH)")
    ap.add_argument("--batch-size", type=int, default=None,
                    help="solutes of a batch workload (256 small / 32 medium)")
    ap.add_argument("--batch-threads", type=int, default=8)
    ap.add_argument("--anderson-m", type=int, default=None,
                    help="history length of Anderson(m) mixing (0: the reference's AN2)")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"],
                    help="fp64: the reference's precision; mixed: FP32 bulk + FP64 sums and "
                         "FP64 steps; both: time both (value = FP64); the small-solute batch "
                         "runs mixed contexts with mixed, FP64 ones otherwise")
    ap.add_argument("--beta", type=float, default=0.6,
                    help="mixing coefficient (AN2's m; beta of Anderson(m))")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--maxstep", type=int, default=10000,
                    help="iteration cap (profiling runs only; the metric needs convergence)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device copy of GIB GiB five times (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs one solve per GPU and has no "
              "multi-GPU path; run with --gpus 1", flush=True)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_calibration > 0:
        n = int(args.copy_calibration * (1 << 30)) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 8, "launches": 5}), flush=True)
        return

    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "batch-medium":
        bench_medium(args)
        return
    rism, grid, box, solv, sol, xvk, desc = workload(args.workload)
    solu = {"usig": sol["sig"], "ueps": sol["eps"], "uq": sol["q"], "uxyz": sol["xyz"]}
    log(f"[bench] workload {args.workload}: grid {grid}, natu {len(sol['sig'])}, "
        f"natv {solv['natv']}")
    am = args.anderson_m if args.anderson_m is not None else 0
    clos = closure_code(rism, args, "PSE3" if args.workload == "cfg3" else "KH")
    g = rism.RISM3D(grid, box, solv, solu, clos, 2, args.beta, 0.2, anderson_m=am)
    for i in range(args.warmup):
        th, tm = g.run(maxstep=args.maxstep, tol=1e-6)
        log(f"[bench] warmup {i}: {tm}")
    g.kernel_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = []
    for i in range(args.steps):
        th, tm = g.run(maxstep=args.maxstep, tol=1e-6)
        results.append((th, tm))
        log(f"[bench] step {i}: {tm} SFE_SC {th['SFE_SC']:.5f}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ks = g.kernel_stats()
    g.kernel_timing(False)
    N = grid[0] * grid[1] * grid[2]
    # h of the FP64 solve for the rms(dh) of the mixed one (4.3 GB of host
    # memory at 512^3 and 4 sites)
    h64 = g.fields(rism.FIELD_H) if args.precision in ("mixed", "both") else None
    g.close()
    th, tm = results[-1]
    if tm["status"] != rism.STATUS_CONVERGED:
        log(f"[bench] WARNING: solve did not converge (status {tm['status']})")
    ms_per_step = 1000.0 * wall / args.steps
    natv = solv["natv"]
    iters = tm["steps"]

    # roofline of the dominant hot-loop kernel
    ab = algorithmic_bytes(natv, N, am if am > 0 else 2)
    hot = {k: v for k, v in ks.items() if k in ab and v[0] > 0}
    dom = max(hot, key=lambda k: hot[k][1])
    launches, total_ms = hot[dom]
    avg_s = total_ms / launches / 1000.0
    achieved = ab[dom] / avg_s / 1e9
    per_kernel = {k: {"launches": v[0], "avg_ms": v[1] / max(v[0], 1),
                      "GB/s": (ab[k] / (v[1] / v[0] / 1000.0) / 1e9) if k in ab and v[0] else None}
                  for k, v in ks.items() if v[0] > 0}
    iter_bytes = sum(ab.values())
    iter_ms = tm["iterate_ms"] / max(iters, 1)
    # the fused lower bound of one iteration, each 3D transform counted as
    # one HBM round trip: N [(12 + 3 m) n + 2] w bytes, m history pairs (2
    # for AN2), w bytes per bulk value
    mh = am if am > 0 else 2

    def fused_bytes(w):
        return N * ((12 + 3 * mh) * natv + 2) * w

    log("[bench] per-kernel:", json.dumps(per_kernel, indent=None))

    mixed = None
    if args.precision in ("mixed", "both"):
        gm = rism.RISM3D(grid, box, solv, solu, clos, 2, args.beta, 0.2, anderson_m=am,
                         precision=1)
        for i in range(args.warmup):
            thm, tmm = gm.run(maxstep=args.maxstep, tol=1e-6)
            log(f"[bench] mixed warmup {i}: {tmm}")
        gm.kernel_timing(True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for i in range(args.steps):
            thm, tmm = gm.run(maxstep=args.maxstep, tol=1e-6)
            log(f"[bench] mixed step {i}: {tmm} SFE_SC {thm['SFE_SC']:.5f}")
        torch.cuda.synchronize()
        wall_m = time.perf_counter() - t1
        ksm = gm.kernel_stats()
        gm.kernel_timing(False)
        rms_dh = None
        if h64 is not None:
            h32 = gm.fields(rism.FIELD_H)
            h32 -= h64
            rms_dh = float(np.sqrt(np.mean(h32 * h32)))
            del h32
        gm.close()
        dmu = thm["SFE_SC"] - th["SFE_SC"]
        mixed = {
            "s_per_solve": wall_m / args.steps,
            "speedup_vs_fp64": (ms_per_step / 1000.0) / (wall_m / args.steps),
            "iterations": tmm["steps"], "iterations_fp32": tmm["steps_fp32"],
            "status": tmm["status"], "setup_ms": tmm["setup_ms"],
            "iterate_ms": tmm["iterate_ms"],
            "SFE_SC": thm["SFE_SC"], "dSFE_SC_J_per_mol": dmu,
            "dSFE_SC_kcal_per_mol": dmu / 4184.0,
            "gate_0.05_kcal_per_mol": abs(dmu) <= 0.05 * 4184.0,
            "dPMV": thm["PMV"] - th["PMV"],
            "rms_dh": rms_dh,
            # FP32 steps at w = 4, FP64 steps at w = 8, over the iteration phase
            "iteration_fused_bound_frac": ((tmm["steps_fp32"] * fused_bytes(4)
                                            + (tmm["steps"] - tmm["steps_fp32"]) * fused_bytes(8))
                                           / (tmm["iterate_ms"] / 1000.0) / (HBM_PEAK_GBS * 1e9)),
            "kernels_avg_ms": {k: v[1] / max(v[0], 1) for k, v in ksm.items() if v[0] > 0},
        }
        log(f"[bench] mixed: {mixed}")

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) ...")
        cpu = cpu_baseline(grid, box, os.path.join(GOLDEN, xvk), sol, iters, natv)
        log(f"[bench] CPU baseline: {cpu}")

    out = {
        "metric": METRIC,
        "value": ms_per_step / 1000.0,
        "unit": "s/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "grid": list(grid), "box": list(box), "natu": int(len(sol["sig"])),
            "natv": int(natv), "closure": args.closure or ("PSE3" if args.workload == "cfg3"
                                                            else "KH"),
            "anderson_m": am, "iterations": iters, "status": tm["status"],
            "setup_ms": tm["setup_ms"], "iterate_ms": tm["iterate_ms"],
            "thermo_ms": tm["thermo_ms"], "ms_per_iteration": iter_ms,
            "iteration_GBps_algorithmic": iter_bytes / (iter_ms / 1000.0) / 1e9,
            "iteration_fused_bound_GB": fused_bytes(8) / 1e9,
            "iteration_fused_bound_frac": fused_bytes(8) / (iter_ms / 1000.0) / (HBM_PEAK_GBS * 1e9),
            "SFE_SC": th["SFE_SC"], "SFE_GF": th["SFE_GF"], "PMV": th["PMV"],
            "kernels": per_kernel,
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (MEASURED_TRAFFIC_CFG2.get(dom)
                        if args.workload == "cfg2" and am == 0 else None),
            "algorithmic_bytes_per_launch": ab[dom],
            "avg_launch_ms": avg_s * 1000.0,
        },
        "cpu_baseline": cpu,
        "mixed": mixed,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
