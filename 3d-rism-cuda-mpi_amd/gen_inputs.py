"""Seeded synthetic solutes for the benchmark and test workloads.

Two generators (numpy PCG64 streams, so a seed fixes the solute exactly):

* random_walk_solute(n, seed, total_charge): atom 0 at the origin, then 1.5 A
  steps in uniform random directions, each kept only if it lies >= 1.2 A from
  every earlier atom (up to 1000 tries per step, else the walk restarts).
  Types uniform over C, N, O, H; charges ~ N(0, 0.3 e) shifted to the total.
* cluster_solute(n, radius, seed, total_charge): random sequential addition,
  uniform in a sphere, >= 1.0 A separation; S with p = 0.01, C/N/O/H with
  p = 0.2475 each; charges ~ N(0, 0.35 e) shifted to the total.

LJ parameters (sigma A, epsilon kcal/mol) are converted to J/mol (x 4184),
the unit of the .inp format.  write_inp() writes the reference's .inp layout
(read_input.cu:32-76).
"""
from __future__ import annotations

import numpy as np

KCAL = 4184.0
TYPES = {  # sigma (A), epsilon (kcal/mol)
    "C": (3.40, 0.086),
    "N": (3.25, 0.17),
    "O": (2.96, 0.21),
    "H": (2.50, 0.015),
    "S": (3.56, 0.25),
}


def _charges(rng, n, sd, total):
    q = rng.normal(0.0, sd, n)
    return q + (total - q.sum()) / n


def _params(names):
    sig = np.array([TYPES[t][0] for t in names])
    eps = np.array([TYPES[t][1] * KCAL for t in names])
    return sig, eps


def random_walk_solute(n, seed, total_charge=0.0, step=1.5, dmin=1.2):
    rng = np.random.default_rng(seed)
    while True:
        xyz = [np.zeros(3)]
        ok = True
        while len(xyz) < n:
            placed = False
            for _ in range(1000):
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                p = xyz[-1] + step * d
                if np.min(np.linalg.norm(np.array(xyz) - p, axis=1)) >= dmin:
                    xyz.append(p)
                    placed = True
                    break
            if not placed:
                ok = False
                break
        if ok:
            break
    names = rng.choice(["C", "N", "O", "H"], size=n)
    sig, eps = _params(names)
    q = _charges(rng, n, 0.3, total_charge)
    return {"sig": sig, "eps": eps, "q": q, "xyz": np.array(xyz)}


def cluster_solute(n, radius, seed, total_charge=0.0, dmin=1.0):
    rng = np.random.default_rng(seed)
    cell = dmin
    grid = {}
    pts = np.zeros((n, 3))
    k = 0
    r2min = dmin * dmin
    while k < n:
        p = rng.uniform(-radius, radius, 3)
        if p @ p > radius * radius:
            continue
        c = tuple(np.floor(p / cell).astype(int))
        clash = False
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((c[0] + dx, c[1] + dy, c[2] + dz), ()):
                        d = pts[j] - p
                        if d @ d < r2min:
                            clash = True
                            break
                    if clash:
                        break
                if clash:
                    break
            if clash:
                break
        if clash:
            continue
        pts[k] = p
        grid.setdefault(c, []).append(k)
        k += 1
    u = rng.uniform(size=n)
    names = np.where(u < 0.01, "S",
                     np.array(["C", "N", "O", "H"])[np.minimum(((u - 0.01) / 0.2475).astype(int), 3)])
    sig, eps = _params(names)
    q = _charges(rng, n, 0.35, total_charge)
    return {"sig": sig, "eps": eps, "q": q, "xyz": pts}


def write_inp(path, solute, xvk, grid, box, closure="KH", tol=1e-6, maxstep=10000,
              an2=(2, 0.6, 0.2), outlist="m", ksave=0):
    with open(path, "w") as f:
        f.write(f" {outlist}    {ksave}  20220301\n {closure}\n {xvk}\n {tol:g}   {maxstep}\n")
        f.write(f" {an2[0]}   {an2[1]}   {an2[2]}\n")
        f.write(f"  {box[0]:g} {box[1]:g} {box[2]:g}\n  {grid[0]} {grid[1]} {grid[2]}\n")
        n = len(solute["sig"])
        f.write(f" {n}\n")
        for i in range(n):
            x, y, z = solute["xyz"][i]
            f.write(f"{solute['sig'][i]:.4f} {solute['eps'][i]:.4f} {solute['q'][i]:.6f} "
                    f"{x:.6f} {y:.6f} {z:.6f}\n")


# The benchmark / test workloads
def cfg2_solute():
    """12877-atom protein-like cluster, R = 30 A, net charge +2 e."""
    return cluster_solute(12877, 30.0, seed=1002, total_charge=2.0)


def cfg1_solute():
    """24-atom random-walk solute, net charge 0."""
    return random_walk_solute(24, seed=1001, total_charge=0.0)


# per-axis grid sizes of the batched small-solute workload (all 5-smooth,
# multiples of 4) and its spacing and padding
BATCH_SIZES = (48, 60, 64, 72, 80, 96)
BATCH_SPACING = 0.5
BATCH_PAD = 12.0


def batch_solute(i, seed0=1004000):
    """Solute i of the batched workload: 8-60 atoms (uniform), random walk,
    net charge -1, 0 or +1 (uniform), from seed seed0 + i."""
    rng = np.random.default_rng(seed0 + i)
    n = int(rng.integers(8, 61))
    q = float(rng.integers(-1, 2))
    return random_walk_solute(n, seed=seed0 + i, total_charge=q)


def batch_grid(xyz, sizes=BATCH_SIZES, spacing=BATCH_SPACING, pad=BATCH_PAD):
    """Per axis the smallest size whose box (size * spacing) holds the solute
    extent plus `pad` on both sides; the largest size when none does."""
    ext = xyz.max(0) - xyz.min(0)
    grid = []
    for e in ext:
        fit = [n for n in sizes if n * spacing >= e + 2 * pad]
        grid.append(fit[0] if fit else sizes[-1])
    return tuple(grid), tuple(n * spacing for n in grid)


def cfg3_solute():
    """3000-atom protein-like cluster, R = 19 A, net charge 0."""
    return cluster_solute(3000, 19.0, seed=1003, total_charge=0.0)


# per-axis grid sizes of the batched medium-size workload
MEDIUM_SIZES = (128, 144, 160, 180, 192)


def medium_solute(i, seed0=1005000):
    """Solute i of the batched medium-size workload: N log-uniform on
    [500, 4000], a cluster of radius (3N / (4 pi 0.1))^(1/3) (0.1 atoms per
    A^3), net charge 0, from seed seed0 + i."""
    rng = np.random.default_rng(seed0 + i)
    n = int(round(np.exp(rng.uniform(np.log(500.0), np.log(4000.0)))))
    radius = (3.0 * n / (4.0 * np.pi * 0.1)) ** (1.0 / 3.0)
    return cluster_solute(n, radius, seed=seed0 + i, total_charge=0.0)
