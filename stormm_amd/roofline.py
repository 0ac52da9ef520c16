"""Algorithmic work of one batched energy / force evaluation, for roofline figures.

Operation counts per unit follow the arithmetic of the reference's formulas as written in the
evaluation kernel (fused multiply-add = 2, any other floating-point operation = 1, a
transcendental or reciprocal / square-root instruction = 1).  Where the Born-radius integral
branches, the far-field (Taylor) branch is counted: it is the cheapest branch and the most common
one in the large systems, so the figures are a lower bound on the arithmetic performed.

Per unordered atom pair (P = N (N - 1) / 2 per system):
  Born-radius integral, both directions ................ 42
  Coulomb + Lennard-Jones + GB pair energy and force .. 74   (energy only: 42)
  GB radius-derivative forces ......................... 56   (not run for energy only)
Per valence term (double precision): 60 on average over bonds (25), angles (55) and cosine
dihedrals with their 1:4 pair (165).
Per atom (Born radius, self energy, chain rule, double precision): 40.

Bytes per evaluation that must cross HBM at least once: fixed-precision coordinates (24 B per
atom), per-atom parameters (72 B), forces read and written (48 B), valence term records and
parameters (16 + 16 B per term, 48 B for dihedrals with their 1:4 data) and tile tasks (16 B per
16 x 16 tile plus 64 B per exclusion mask).
"""
from __future__ import annotations

FLOP_PAIR_FORCE = 42 + 74 + 56
FLOP_PAIR_ENERGY = 42 + 42
FLOP_TERM = 60
FLOP_ATOM = 40

# MI355X datasheet peaks.
PEAK_FP32_TFLOPS = 157.3       # dense FP32 (vector and matrix share this figure)
PEAK_FP64_TFLOPS = 78.6        # dense FP64 vector
PEAK_HBM_GBPS = 8000.0


def eval_flops(pairs: int, terms: int, atoms: int, forces: bool = True) -> float:
    per_pair = FLOP_PAIR_FORCE if forces else FLOP_PAIR_ENERGY
    return float(pairs) * per_pair + float(terms) * FLOP_TERM + float(atoms) * FLOP_ATOM


def eval_bytes(atoms: int, terms: int, tasks: int, forces: bool = True) -> float:
    per_atom = 24 + 72 + (48 if forces else 0)
    return float(atoms) * per_atom + float(terms) * 40 + float(tasks) * 16
