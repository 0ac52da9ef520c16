"""Seeded synthetic branched-chain molecules for the benchmark and parity workloads.

Each molecule is a tree of atoms grown one atom at a time: a new atom bonds to an existing atom
with fewer than four bonds, sits at the bond's equilibrium length, makes the equilibrium angle
with its parent's parent, and takes a uniformly random torsion (re-drawn when it would land
within 2.2 A of an atom more than two bonds away).  Coordinates then get N(0, 0.05 A) jitter.

Parameters follow the workload definitions:
  bonds       k ~ U[200, 400] kcal/mol/A^2, r0 ~ U[1.0, 1.55] A
  angles      k ~ U[40, 80] kcal/mol/rad^2, theta0 ~ U[105, 120] deg
  dihedrals   one cosine term per proper torsion, n in {1, 2, 3}, barrier ~ U[0, 2] kcal/mol,
              phase 0 or 180 deg
  LJ          one type per atom, sigma ~ U[2.0, 3.8] A, epsilon ~ U[0.02, 0.2] kcal/mol,
              Lorentz-Berthelot tables
  charges     N(0, 0.3) e shifted to zero net charge and placed on the 2^-23 e grid with an
              exact zero sum (so the reference's charge smoothing leaves them unchanged)
  1:4         electrostatics / 1.2, Lennard-Jones / 2
  GB          intrinsic radii ~ U[1.2, 2.0] A, screening ~ U[0.7, 0.9], dielectric 78.5
"""
from __future__ import annotations

import math
import multiprocessing
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .topology import (AMBER_COULOMB, AMBER_SCEE, AMBER_SCNB, IGB_NONE, IGB_OBC2, Topology)

CHARGE_GRID = 2.0 ** -23


def _grid_charges(rng: np.random.Generator, n: int) -> np.ndarray:
    q = rng.normal(0.0, 0.3, n)
    q -= q.mean()
    iq = np.rint(q / CHARGE_GRID).astype(np.int64)
    resid = int(iq.sum())
    # Remove the residual one grid step at a time from the largest-magnitude charges.
    order = np.argsort(-np.abs(iq), kind="stable")
    k = 0
    while resid != 0:
        step = 1 if resid < 0 else -1
        iq[order[k % n]] += step
        resid += step
        k += 1
    return iq.astype(np.float64) * CHARGE_GRID


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _place(parent: np.ndarray, grand: np.ndarray, great: np.ndarray, r0: float, theta: float,
           phi: float) -> np.ndarray:
    """Position at distance r0 from parent, angle theta to grand, torsion phi to great."""
    px, py, pz = float(parent[0]), float(parent[1]), float(parent[2])
    bc = (px - float(grand[0]), py - float(grand[1]), pz - float(grand[2]))
    lb = math.sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2])
    bc = (bc[0] / lb, bc[1] / lb, bc[2] / lb)
    ab = (float(grand[0]) - float(great[0]), float(grand[1]) - float(great[1]),
          float(grand[2]) - float(great[2]))
    nrm = _cross(ab, bc)
    ln = math.sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2])
    if ln < 1.0e-8:
        trial = (1.0, 0.0, 0.0) if abs(bc[0]) < 0.9 else (0.0, 1.0, 0.0)
        nrm = _cross(trial, bc)
        ln = math.sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2])
    nrm = (nrm[0] / ln, nrm[1] / ln, nrm[2] / ln)
    m = _cross(nrm, bc)
    d0 = -r0 * math.cos(theta)
    d1 = r0 * math.sin(theta) * math.cos(phi)
    d2 = r0 * math.sin(theta) * math.sin(phi)
    return np.array([px + d0 * bc[0] + d1 * m[0] + d2 * nrm[0],
                     py + d0 * bc[1] + d1 * m[1] + d2 * nrm[1],
                     pz + d0 * bc[2] + d1 * m[2] + d2 * nrm[2]])


def branched_molecule(natom: int, seed: int, gb: bool = True,
                      jitter: float = 0.05) -> Tuple[Topology, np.ndarray]:
    """Generate one molecule of `natom` atoms; returns (topology, coordinates (N, 3))."""
    if natom < 4:
        raise ValueError("a branched chain needs at least 4 atoms")
    rng = np.random.default_rng(seed)
    parent = [-1]
    nbond = [0] * natom
    bond_r0: List[float] = []
    bond_k: List[float] = []
    bonds: List[Tuple[int, int]] = []
    theta_of = {}  # (grand, parent, child) -> theta0 used when placing
    xyz = np.zeros((natom, 3))
    # Seed three atoms in a bent chain.
    r01 = rng.uniform(1.0, 1.55)
    r12 = rng.uniform(1.0, 1.55)
    th = math.radians(rng.uniform(105.0, 120.0))
    xyz[1] = [r01, 0.0, 0.0]
    xyz[2] = xyz[1] + r12 * np.array([-math.cos(th), math.sin(th), 0.0])
    parent += [0, 1]
    bonds += [(0, 1), (1, 2)]
    bond_r0 += [r01, r12]
    bond_k += [rng.uniform(200.0, 400.0), rng.uniform(200.0, 400.0)]
    nbond[0], nbond[1], nbond[2] = 1, 2, 1
    theta_of[(0, 1, 2)] = th
    adj: List[List[int]] = [[1], [0, 2], [1]] + [[] for _ in range(natom - 3)]

    for new in range(3, natom):
        placed = False
        for attempt in range(60):
            # Prefer recent atoms (chain growth) but allow branching anywhere.
            lo = max(1, new - 6) if attempt < 40 else 1
            cands = [a for a in range(lo, new) if nbond[a] < 4 and parent[a] >= 0]
            if not cands:
                cands = [a for a in range(1, new) if nbond[a] < 4 and parent[a] >= 0]
            par = int(rng.choice(cands))
            grand = parent[par]
            # The torsion reference is the grandparent's own parent, or an arbitrary direction
            # at the root (the torsion is random either way).
            great_pos = xyz[parent[grand]] if parent[grand] >= 0 else \
                xyz[grand] + np.array([0.0, 0.0, 1.0])
            r0 = rng.uniform(1.0, 1.55)
            theta = math.radians(rng.uniform(105.0, 120.0))
            phi = rng.uniform(-math.pi, math.pi)
            pos = _place(xyz[par], xyz[grand], great_pos, r0, theta, phi)
            # Clash rules: siblings (1:3 through the parent) at least 1.8 A apart, every atom
            # three or more bonds away at least 2.2 A away.
            dist = np.sqrt(np.sum((xyz[:new] - pos) ** 2, axis=1))
            limit = np.full(new, 2.2)
            limit[par] = 0.0
            limit[grand] = 0.0
            for a in adj[par]:
                if a != grand:
                    limit[a] = 1.8
            if np.all(dist >= limit):
                placed = True
                break
        if not placed:
            # Accept the last trial; the minimizer handles the strain.
            pass
        xyz[new] = pos
        parent.append(par)
        bonds.append((par, new))
        bond_r0.append(r0)
        bond_k.append(rng.uniform(200.0, 400.0))
        nbond[par] += 1
        nbond[new] = 1
        adj[par].append(new)
        adj[new].append(par)
        theta_of[(grand, par, new)] = theta
    xyz += rng.normal(0.0, jitter, xyz.shape)

    # Angles: every pair of neighbours around a centre.
    angles, angle_k, angle_theta = [], [], []
    for j in range(natom):
        nb = sorted(adj[j])
        for a in range(len(nb)):
            for c in range(a + 1, len(nb)):
                i, k = nb[a], nb[c]
                angles.append((i, j, k))
                angle_k.append(rng.uniform(40.0, 80.0))
                th0 = theta_of.get((i, j, k), theta_of.get((k, j, i), None))
                if th0 is None:
                    th0 = math.radians(rng.uniform(105.0, 120.0))
                angle_theta.append(th0)
    # Proper dihedrals: every i-j-k-l path, one cosine term each, each carrying its 1:4 pair.
    dihedrals, amp, phase, freq = [], [], [], []
    for (j, k) in bonds:
        for i in adj[j]:
            if i == k:
                continue
            for l in adj[k]:
                if l == j or l == i:
                    continue
                dihedrals.append((i, j, k, l))
                freq.append(float(rng.integers(1, 4)))
                amp.append(rng.uniform(0.0, 2.0))
                phase.append(0.0 if rng.random() < 0.5 else math.pi)
    nd = len(dihedrals)
    sigma = rng.uniform(2.0, 3.8, natom)
    eps = rng.uniform(0.02, 0.2, natom)
    q = _grid_charges(rng, natom)
    radii = rng.uniform(1.2, 2.0, natom)
    screen = rng.uniform(0.7, 0.9, natom)
    topo = Topology(
        natom=natom, charge=q, lj_idx=np.arange(natom, dtype=np.int32), n_lj_types=natom,
        bonds=np.array(bonds, dtype=np.int32).reshape(-1, 2),
        bond_k=np.array(bond_k), bond_r0=np.array(bond_r0),
        angles=np.array(angles, dtype=np.int32).reshape(-1, 3),
        angle_k=np.array(angle_k), angle_theta=np.array(angle_theta),
        dihedrals=np.array(dihedrals, dtype=np.int32).reshape(-1, 4),
        dihe_amp=np.array(amp), dihe_phase=np.array(phase), dihe_freq=np.array(freq),
        dihe_improper=np.zeros(nd, dtype=np.int32), dihe_scee=np.full(nd, AMBER_SCEE),
        dihe_scnb=np.full(nd, AMBER_SCNB), pb_radii=radii, gb_screen=screen,
        igb=IGB_OBC2 if gb else IGB_NONE, dielectric=78.5, salt=0.0, coulomb=AMBER_COULOMB,
        name=f"branched_{natom}_{seed}", lj_sigma=sigma, lj_epsilon=eps)
    return topo, xyz


# ----------------------------------------------------------------------------------------------
# Workloads
# ----------------------------------------------------------------------------------------------
def config0() -> Tuple[List[Topology], List[np.ndarray]]:
    """32 copies of one seeded 48-atom molecule, vacuum."""
    t, x = branched_molecule(48, seed=20240, gb=False)
    return [t] * 32, [x.copy() for _ in range(32)]


def _make_one(args):
    n, mseed, gb, perturb, pseed = args
    t, x = branched_molecule(int(n), seed=mseed, gb=gb)
    if perturb > 0.0:
        x = x + np.random.default_rng(pseed).normal(0.0, perturb, x.shape)
    return t, x


def generate(sizes: Sequence[int], seeds: Sequence[int], gb: bool = True, perturb: float = 0.0,
             perturb_seeds: Optional[Sequence[int]] = None,
             workers: Optional[int] = None) -> Tuple[List[Topology], List[np.ndarray]]:
    """Generate molecules of the given sizes and seeds, in parallel for large batches."""
    if perturb_seeds is None:
        perturb_seeds = [s * 7 + 1 for s in seeds]
    jobs = [(int(n), int(sd), gb, perturb, int(ps)) for n, sd, ps in
            zip(sizes, seeds, perturb_seeds)]
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if len(jobs) >= 64 and workers > 1:
        ctx = multiprocessing.get_context("fork")
        with ctx.Pool(workers) as pool:
            out = pool.map(_make_one, jobs, chunksize=max(1, len(jobs) // (workers * 8)))
    else:
        out = [_make_one(j) for j in jobs]
    return [o[0] for o in out], [o[1] for o in out]


def uniform_batch(nsys: int, lo: int, hi: int, seed: int, gb: bool = True,
                  perturb: float = 0.0) -> Tuple[List[Topology], List[np.ndarray]]:
    """nsys systems with atom counts ~ U[lo, hi] and per-system seeds."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, nsys)
    seeds = [seed * 100003 + s for s in range(nsys)]
    pseeds = [seed * 7 + s for s in range(nsys)]
    return generate(sizes, seeds, gb=gb, perturb=perturb, perturb_seeds=pseeds)


def config1(nsys: int = 1024, seed: int = 1) -> Tuple[List[Topology], List[np.ndarray]]:
    """Heterogeneous systems, atom counts ~ U[24, 256], OBC-II GB."""
    return uniform_batch(nsys, 24, 256, seed, gb=True)


def config2(nsys: int = 4096, seed: int = 2) -> Tuple[List[Topology], List[np.ndarray]]:
    """Minimization batch, atom counts ~ U[16, 128], GB, starting points perturbed N(0, 0.3 A)."""
    return uniform_batch(nsys, 16, 128, seed, gb=True, perturb=0.3)


def config4(nsys: int = 16384, seed: int = 4) -> Tuple[List[Topology], List[np.ndarray]]:
    """Large batch, atom counts ~ U[32, 96], GB."""
    return uniform_batch(nsys, 32, 96, seed, gb=True)


def log_uniform_sizes(rng: np.random.Generator, count: int, lo: int = 16,
                      hi: int = 512) -> np.ndarray:
    return np.exp(rng.uniform(math.log(lo), math.log(hi), count)).astype(np.int64).clip(lo, hi)
