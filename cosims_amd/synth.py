"""Synthetic ions in the reference's MFJ input format.

Molecules are written as MFJ files (name / 1 / N / ang / calc / 1.0, then
rows ``x y z intMass q``) so they load through the reference reader
(src/molecule2.cpp:442-647) and the engine's own reader alike.

Recipe: C/H/N/O/P in the ratio 30:50:10:8:2 rounded by largest remainder;
positions by seeded random sequential addition, uniform in a ball of radius R
with a minimum pair separation; charges q_i ~ N(0, sigma_q) shifted so that
they sum to the net charge.
"""
from __future__ import annotations

import math

import numpy as np

TYPE_INTMASS = (12, 1, 14, 16, 31)      # C, H, N, O, P
TYPE_RATIO = (30, 50, 10, 8, 2)


def type_counts(n: int, ratio=TYPE_RATIO):
    tot = float(sum(ratio))
    exact = [n * r / tot for r in ratio]
    counts = [int(math.floor(e)) for e in exact]
    rem = n - sum(counts)
    order = sorted(range(len(ratio)), key=lambda i: (-(exact[i] - counts[i]), i))
    for i in order[:rem]:
        counts[i] += 1
    return counts


def synthetic_ion(n: int, radius: float, sigma_q: float, net_charge: float, seed: int,
                  min_sep: float = 1.0):
    """Return dict(xyz[n,3], intmass[n], charge[n]) for a seeded synthetic ion."""
    rng = np.random.default_rng(seed)
    counts = type_counts(n)
    intmass = np.repeat(np.array(TYPE_INTMASS), counts)
    rng.shuffle(intmass)
    cell = min_sep
    grid: dict = {}
    pts = np.empty((n, 3))
    k = 0
    sep2 = min_sep * min_sep
    attempts = 0
    while k < n:
        attempts += 1
        if attempts > 200 * n + 10000:
            raise RuntimeError("packing failed; radius too small for n atoms")
        p = rng.uniform(-radius, radius, 3)
        if p @ p > radius * radius:
            continue
        c = tuple(np.floor(p / cell).astype(int))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((c[0] + dx, c[1] + dy, c[2] + dz), ()):
                        d = pts[j] - p
                        if d @ d < sep2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        pts[k] = p
        grid.setdefault(c, []).append(k)
        k += 1
    q = rng.normal(0.0, sigma_q, n)
    q += (net_charge - q.sum()) / n
    return dict(xyz=pts, intmass=intmass, charge=q)


def write_mfj(path: str, name: str, ion: dict):
    xyz, im, q = ion["xyz"], ion["intmass"], ion["charge"]
    with open(path, "w") as f:
        f.write(f"{name}\n1\n{len(im)}\nang\ncalc\n1.0\n")
        for i in range(len(im)):
            f.write(f"{xyz[i, 0]:12.6f} {xyz[i, 1]:12.6f} {xyz[i, 2]:12.6f} {int(im[i]):4d} "
                    f"{q[i]:12.8f}\n")


# Workloads (atom count, radius A, sigma_q e, net charge e, molecule seed).
CONFIG_MOLECULES = {
    "ion128": dict(n=128, radius=9.0, sigma_q=0.3, net_charge=1.0, seed=1001),
    "ion2000": dict(n=2000, radius=18.0, sigma_q=0.3, net_charge=8.0, seed=2001),
    "ion20000": dict(n=20000, radius=40.0, sigma_q=0.4, net_charge=25.0, seed=4001),
}


def config_ion(key: str):
    c = CONFIG_MOLECULES[key]
    return synthetic_ion(c["n"], c["radius"], c["sigma_q"], c["net_charge"], c["seed"])


def uneven_conformers(count=256, n_min=50, n_max=5000, seed=3000):
    """The batched uneven workload: atom counts log-uniform in [n_min, n_max],
    radius 18 A x (N/2000)^(1/3), net charge round(N/250)+1 e, sigma_q 0.3 e.
    Returns a list of (N, ion dict)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        ion = synthetic_ion(n, 18.0 * (n / 2000.0) ** (1 / 3), 0.3, round(n / 250) + 1, seed + 1 + k)
        out.append((n, ion))
    return out


def serving_stream(count=40, rate=20.0, n_min=100, n_max=20000, seed=5000):
    """The serving workload: Poisson arrivals at `rate` per second; atom counts
    log-uniform in [n_min, n_max], radius 18 A x (N/2000)^(1/3), net charge
    round(N/800)+1 e, sigma_q 0.3 e, He temperature uniform in 200-400 K.
    Returns a list of dict(t (s after the first arrival), n, temp, ion)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        if k:
            t += rng.exponential(1.0 / rate)
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        temp = float(rng.uniform(200.0, 400.0))
        ion = synthetic_ion(n, 18.0 * (n / 2000.0) ** (1 / 3), 0.3, round(n / 800) + 1, seed + 1 + k)
        out.append(dict(t=t, n=n, temp=temp, ion=ion))
    return out
