"""Flat-bottom restraints: records in the layout of the C-ABI ``stormm_restraint`` and the
builders of src/Restraints/restraint_builder.h.

A restraint holds atom 0 to a target point (positional), atoms 0-1 (distance), 0-1-2 (angle) or
0-1-2-3 (dihedral) within a flat-bottomed well r2..r3, harmonic with stiffness k2 below r2 down
to r1 and k3 above r3 up to r4, linear beyond (src/Restraints/restraint_util.tpp:8-45).  Its
parameters slide linearly from the initial to the final set between init_step and final_step
(restraint_util.tpp:47-77).  A system's restraints are given to ``AtomGraphSynthesis`` as one
record array per system (atom indices local to that system) and feed the RESTRAINT ScoreCard
slot.
"""
from dataclasses import dataclass
from typing import Iterable, List, Optional, Sequence

import numpy as np

POSITION = 1
DISTANCE = 2
ANGLE = 3
DIHEDRAL = 4

# The C struct stormm_restraint: int32 kind, atoms[4], init_step, final_step, then doubles.
RESTRAINT_DTYPE = np.dtype([
    ("kind", np.int32), ("atoms", np.int32, (4,)),
    ("init_step", np.int32), ("final_step", np.int32),
    ("init_k", np.float64, (2,)), ("final_k", np.float64, (2,)),
    ("init_r", np.float64, (4,)), ("final_r", np.float64, (4,)),
    ("init_target", np.float64, (3,)), ("final_target", np.float64, (3,)),
], align=True)
assert RESTRAINT_DTYPE.itemsize == 176


@dataclass
class FlatBottomPlan:
    """One state of a restraint (src/Restraints/restraint_builder.h:20-40,
    restraint_builder.cpp:29-56).  The default has k2 = 0, k3 = 0.3 and every bound at 0."""
    k2: float = 0.0
    k3: float = 0.3
    r1: float = 0.0
    r2: float = 0.0
    r3: float = 0.0
    r4: float = 0.0
    activation_step: int = 0

    @classmethod
    def harmonic(cls, k: float, r: float) -> "FlatBottomPlan":
        """FlatBottomPlan(k, r): a harmonic well at r, linear 10 units away on either side."""
        return cls(k, k, r - 10.0, r, r, r + 10.0, 0)

    @classmethod
    def well(cls, k: float, r2: float, r3: float, k3: Optional[float] = None) -> "FlatBottomPlan":
        """FlatBottomPlan(k, r2, r3) and FlatBottomPlan(k2, k3, r2, r3): flat between r2 and r3,
        linear 10 units outside."""
        return cls(k, k if k3 is None else k3, r2 - 10.0, r2, r3, r3 + 10.0, 0)


def empty(n: int = 0) -> np.ndarray:
    return np.zeros(n, dtype=RESTRAINT_DTYPE)


def check_displacement_limits(kind: int, r) -> tuple:
    """Validated and adjusted bounds (r1, r2, r3, r4) of one restraint state, as
    BoundedRestraint::checkDisplacementLimits (src/Restraints/bounded_restraint.cpp:667-719):
    the bounds must not decrease; an angle's flat bottom lies in [0, pi) and its r1, r4 are
    clipped to [0, pi]; a dihedral's flat bottom is narrower than 2 pi, its r1 .. r4 span at
    most 2 pi, and all four shift together so that the flat bottom's midpoint lies in
    [0, 2 pi)."""
    r1, r2, r3, r4 = (float(v) for v in r)
    if r1 > r2 or r2 > r3 or r3 > r4:
        raise ValueError(f"restraint bounds must not decrease: {r1}, {r2}, {r3}, {r4}")
    pi = np.pi
    twopi = 2.0 * np.pi
    if kind == ANGLE and (r2 < 0.0 or r3 >= pi):
        raise ValueError(f"angle restraints need r2 and r3 in [0, pi): r2 = {r2}, r3 = {r3}")
    if kind == DIHEDRAL and r3 - r2 >= twopi:
        raise ValueError(f"dihedral restraints need r3 - r2 < 2 pi: r2 = {r2}, r3 = {r3}")
    if kind == ANGLE:
        if r1 < 0.0:
            r1 = 0.0
        elif r4 >= pi:
            r4 = pi
    elif kind == DIHEDRAL and r4 - r1 > twopi:
        r1 = r2 - (0.5 * (twopi - (r3 - r2)))
        r4 = r3 + (0.5 * (twopi - (r3 - r2)))
    if kind == DIHEDRAL:
        mid = 0.5 * (r2 + r3)
        frac = mid / twopi
        shift = ((frac - np.floor(frac)) * twopi) - mid
        r1, r2, r3, r4 = r1 + shift, r2 + shift, r3 + shift, r4 + shift
    return r1, r2, r3, r4


def _record(kind: int, atoms: Sequence[int], init: FlatBottomPlan, final: FlatBottomPlan,
            init_target=(0.0, 0.0, 0.0), final_target=None) -> np.ndarray:
    r = np.zeros(1, dtype=RESTRAINT_DTYPE)
    r["kind"] = kind
    a = list(atoms) + [0] * (4 - len(atoms))
    r["atoms"][0] = a
    r["init_step"] = init.activation_step
    r["final_step"] = final.activation_step
    r["init_k"][0] = (init.k2, init.k3)
    r["final_k"][0] = (final.k2, final.k3)
    r["init_r"][0] = check_displacement_limits(kind, (init.r1, init.r2, init.r3, init.r4))
    r["final_r"][0] = check_displacement_limits(kind, (final.r1, final.r2, final.r3, final.r4))
    r["init_target"][0] = init_target
    r["final_target"][0] = init_target if final_target is None else final_target
    return r


def apply_distance_restraint(atom_i: int, atom_j: int, fb_init: FlatBottomPlan,
                             fb_final: Optional[FlatBottomPlan] = None) -> np.ndarray:
    """applyDistanceRestraint, restraint_builder.cpp:59-72."""
    return _record(DISTANCE, (atom_i, atom_j), fb_init, fb_final or fb_init)


def apply_angle_restraint(atom_i: int, atom_j: int, atom_k: int, fb_init: FlatBottomPlan,
                          fb_final: Optional[FlatBottomPlan] = None) -> np.ndarray:
    """applyAngleRestraint, restraint_builder.cpp:75-88 (bounds in radians)."""
    return _record(ANGLE, (atom_i, atom_j, atom_k), fb_init, fb_final or fb_init)


def apply_dihedral_restraint(atom_i: int, atom_j: int, atom_k: int, atom_l: int,
                             fb_init: FlatBottomPlan,
                             fb_final: Optional[FlatBottomPlan] = None) -> np.ndarray:
    """applyDihedralRestraint, restraint_builder.cpp:91-104 (bounds in radians)."""
    return _record(DIHEDRAL, (atom_i, atom_j, atom_k, atom_l), fb_init, fb_final or fb_init)


def apply_positional_restraint(atom: int, target, fb_init: FlatBottomPlan,
                               fb_final: Optional[FlatBottomPlan] = None,
                               final_target=None) -> np.ndarray:
    """One positional restraint (BoundedRestraint, src/Restraints/bounded_restraint.h)."""
    return _record(POSITION, (atom,), fb_init, fb_final or fb_init,
                   tuple(float(v) for v in target),
                   None if final_target is None else tuple(float(v) for v in final_target))


def apply_positional_restraints(ref_xyz: np.ndarray, atoms: Iterable[int],
                                displacement_penalty: float, displacement_onset: float = 0.0,
                                displacement_plateau: float = 16.0,
                                proximity_penalty: float = 0.0, proximity_onset: float = 0.0,
                                proximity_plateau: float = 0.0) -> np.ndarray:
    """applyPositionalRestraints, restraint_builder.cpp:155-170: each listed atom held near its
    reference position, k2 = proximity_penalty, k3 = displacement_penalty,
    (r1, r2, r3, r4) = (proximity_plateau, proximity_onset, displacement_onset,
    displacement_plateau)."""
    ref = np.asarray(ref_xyz, dtype=np.float64).reshape(-1, 3)
    plan = FlatBottomPlan(proximity_penalty, displacement_penalty, proximity_plateau,
                          proximity_onset, displacement_onset, displacement_plateau, 0)
    recs = [apply_positional_restraint(a, ref[a], plan) for a in atoms]
    return np.concatenate(recs) if recs else empty()


def apply_distance_restraints(topo, ref_xyz: np.ndarray, atoms: Sequence[int], penalty: float,
                              flat_bottom_half_width: float, cutoff: float) -> np.ndarray:
    """applyDistanceRestraints, restraint_builder.cpp:193-244: every pair of listed atoms that
    is not a 1:2 or 1:3 pair and lies closer than cutoff in the reference coordinates is held
    within flat_bottom_half_width of its reference distance (r1 = 0, r4 = r3 + 1000)."""
    ref = np.asarray(ref_xyz, dtype=np.float64).reshape(-1, 3)
    near = set()
    for i, j in np.asarray(topo.bonds).reshape(-1, 2):
        near.add((min(i, j), max(i, j)))
    for i, _, k in np.asarray(topo.angles).reshape(-1, 3):
        near.add((min(i, k), max(i, k)))
    atoms = list(atoms)
    recs: List[np.ndarray] = []
    for p, ai in enumerate(atoms):
        for aj in atoms[p + 1:]:
            if (min(ai, aj), max(ai, aj)) in near:
                continue
            target = float(np.linalg.norm(ref[aj] - ref[ai]))
            if target < cutoff:
                r2 = max(0.0, target - flat_bottom_half_width)
                r3 = target + flat_bottom_half_width
                plan = FlatBottomPlan(penalty, penalty, 0.0, r2, r3, r3 + 1000.0, 0)
                recs.append(apply_distance_restraint(ai, aj, plan))
    return np.concatenate(recs) if recs else empty()


def concatenate(parts: Iterable[np.ndarray]) -> np.ndarray:
    parts = [np.asarray(p, dtype=RESTRAINT_DTYPE).reshape(-1) for p in parts]
    return np.concatenate(parts) if parts else empty()
