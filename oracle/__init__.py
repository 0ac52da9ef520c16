"""Python bindings to the CPU double-precision oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the stormm_amd HIP path: only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py may import this package.  Nothing in
stormm_amd imports it, and the product path raises if its own HIP library is missing instead of
falling back to this code.

The C sources restate the reference's CPU evaluation (src/Potential/nonbonded_potential.tpp,
src/Potential/valence_potential.tpp, src/MolecularMechanics/mm_evaluation.tpp) and the GPU
minimizer's conjugate-gradient / line-search semantics (src/Math/hpc_reduction.cu,
src/Math/conjugate_gradient.cui, src/MolecularMechanics/line_movement.cui).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_SOURCES = [os.path.join(_HERE, "stormm_oracle.c")]

SLOT_NAMES = ["BOND", "ANGLE", "PROPER_DIHEDRAL", "IMPROPER_DIHEDRAL", "UREY_BRADLEY",
              "CHARMM_IMPROPER", "CMAP", "VDW", "VDW_ONE_FOUR", "ELECTROSTATIC", "ELEC_ONE_FOUR",
              "GENERALIZED_BORN", "RESTRAINT"]
SLOT = {name: k for k, name in enumerate(SLOT_NAMES)}


# Optimised for the x86-64-v3 level (AVX2 + FMA), which every current x86 server core has, so
# that the library built here also runs on the host of the GPU machine; floating-point
# contraction stays off so that the arithmetic is exactly the source's on any machine.
CFLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11",
          "-D_GNU_SOURCE"]


def build(verbose: bool = False) -> str:
    """Compile the oracle with gcc into oracle/liboracle.so."""
    cmd = ["gcc"] + CFLAGS + ["-o", _LIB_PATH] + _SOURCES + ["-lm", "-lpthread"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    return _LIB_PATH


class OrcSystem(ctypes.Structure):
    _fields_ = [
        ("natom", ctypes.c_int), ("coulomb", ctypes.c_double),
        ("charge", ctypes.c_void_p),
        ("n_lj_types", ctypes.c_int), ("lj_idx", ctypes.c_void_p),
        ("lja", ctypes.c_void_p), ("ljb", ctypes.c_void_p),
        ("lja14", ctypes.c_void_p), ("ljb14", ctypes.c_void_p),
        ("excl", ctypes.c_void_p),
        ("nbond", ctypes.c_int), ("bond_ij", ctypes.c_void_p),
        ("bond_keq", ctypes.c_void_p), ("bond_leq", ctypes.c_void_p),
        ("nangl", ctypes.c_int), ("angl_ijk", ctypes.c_void_p),
        ("angl_keq", ctypes.c_void_p), ("angl_theta", ctypes.c_void_p),
        ("ndihe", ctypes.c_int), ("dihe_ijkl", ctypes.c_void_p),
        ("dihe_amp", ctypes.c_void_p), ("dihe_phase", ctypes.c_void_p),
        ("dihe_freq", ctypes.c_void_p), ("dihe_improper", ctypes.c_void_p),
        ("dihe_scee", ctypes.c_void_p), ("dihe_scnb", ctypes.c_void_p),
        ("ninfr14", ctypes.c_int), ("infr14_il", ctypes.c_void_p),
        ("infr14_scee", ctypes.c_void_p), ("infr14_scnb", ctypes.c_void_p),
        ("igb", ctypes.c_int), ("dielectric", ctypes.c_double), ("kappa", ctypes.c_double),
        ("gb_offset", ctypes.c_double),
        ("pb_radii", ctypes.c_void_p), ("gb_screen", ctypes.c_void_p),
        ("gb_alpha", ctypes.c_void_p), ("gb_beta", ctypes.c_void_p),
        ("gb_gamma", ctypes.c_void_p),
        ("nrst", ctypes.c_int), ("rst", ctypes.c_void_p),
    ]


# orc_restraint (stormm_oracle.h): kind, atoms[4], init_step, final_step, k[2] x 2, r[4] x 2,
# target[3] x 2.
ORC_RESTRAINT = np.dtype([
    ("kind", np.int32), ("atoms", np.int32, (4,)),
    ("init_step", np.int32), ("final_step", np.int32),
    ("init_k", np.float64, (2,)), ("final_k", np.float64, (2,)),
    ("init_r", np.float64, (4,)), ("final_r", np.float64, (4,)),
    ("init_target", np.float64, (3,)), ("final_target", np.float64, (3,)),
], align=True)


class OrcEnergy(ctypes.Structure):
    _fields_ = [("nrg", ctypes.c_double * 32), ("acc", ctypes.c_longlong * 32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        L.orc_eval_mm.argtypes = [ctypes.POINTER(OrcSystem), vp, vp, vp, vp, vp, vp,
                                  ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcEnergy)]
        L.orc_eval_mm.restype = None
        L.orc_eval_mm_clash.argtypes = [ctypes.POINTER(OrcSystem), vp, vp, vp, vp, vp, vp,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.POINTER(OrcEnergy)]
        L.orc_eval_mm_clash.restype = None
        L.orc_eval_mm_ext.argtypes = [ctypes.POINTER(OrcSystem), vp, vp, vp, vp, vp, vp,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, ctypes.POINTER(OrcEnergy)]
        L.orc_eval_mm_ext.restype = None
        L.orc_eval_terms.argtypes = [ctypes.POINTER(OrcSystem), vp, vp, vp, vp, vp, vp,
                                     ctypes.c_uint, ctypes.c_int, ctypes.POINTER(OrcEnergy)]
        L.orc_eval_terms.restype = None
        L.orc_gb.argtypes = [ctypes.POINTER(OrcSystem), vp, vp, vp, vp, vp, vp, ctypes.c_int,
                             ctypes.c_double, ctypes.POINTER(OrcEnergy), vp, vp, vp]
        L.orc_gb.restype = None
        L.orc_eval_batch.argtypes = [ctypes.c_int, ctypes.POINTER(OrcSystem), vp, vp, vp, vp,
                                     vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.orc_eval_batch.restype = ctypes.c_double
        L.orc_cubic_move.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_double), vp]
        L.orc_cubic_move.restype = ctypes.c_double
        L.orc_minimize.argtypes = [ctypes.POINTER(OrcSystem), vp, vp, vp, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, vp]
        L.orc_minimize.restype = None
        L.orc_minimize_record.argtypes = L.orc_minimize.argtypes + [vp]
        L.orc_minimize_record.restype = None
        L.orc_plant_mistake.argtypes = [ctypes.c_int]
        L.orc_plant_mistake.restype = None
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class SystemHandle:
    """Keeps the numpy arrays behind one orc_system alive.  `restraints` is a record array with
    the fields of ORC_RESTRAINT (atom indices local to the system) or None."""

    def __init__(self, topo, restraints=None):
        n = topo.natom
        keep: List[np.ndarray] = []

        def f64(v):
            a = np.ascontiguousarray(v, dtype=np.float64)
            keep.append(a)
            return a

        def i32(v):
            a = np.ascontiguousarray(v, dtype=np.int32)
            keep.append(a)
            return a

        s = OrcSystem()
        s.natom = n
        s.coulomb = topo.coulomb
        s.charge = _ptr(f64(topo.charge))
        s.n_lj_types = topo.n_lj_types
        s.lj_idx = _ptr(i32(topo.lj_idx))
        lja, ljb, lja14, ljb14 = topo.lj_tables()
        s.lja = _ptr(f64(lja))
        s.ljb = _ptr(f64(ljb))
        s.lja14 = _ptr(f64(lja14))
        s.ljb14 = _ptr(f64(ljb14))
        ex = np.ascontiguousarray(topo.exclusion_matrix(), dtype=np.uint8)
        keep.append(ex)
        s.excl = _ptr(ex)
        s.nbond = len(topo.bonds)
        s.bond_ij = _ptr(i32(np.asarray(topo.bonds).reshape(-1)))
        s.bond_keq = _ptr(f64(topo.bond_k))
        s.bond_leq = _ptr(f64(topo.bond_r0))
        s.nangl = len(topo.angles)
        s.angl_ijk = _ptr(i32(np.asarray(topo.angles).reshape(-1)))
        s.angl_keq = _ptr(f64(topo.angle_k))
        s.angl_theta = _ptr(f64(topo.angle_theta))
        s.ndihe = len(topo.dihedrals)
        s.dihe_ijkl = _ptr(i32(np.asarray(topo.dihedrals).reshape(-1)))
        s.dihe_amp = _ptr(f64(topo.dihe_amp))
        s.dihe_phase = _ptr(f64(topo.dihe_phase))
        s.dihe_freq = _ptr(f64(topo.dihe_freq))
        s.dihe_improper = _ptr(i32(topo.dihe_improper))
        s.dihe_scee = _ptr(f64(topo.dihe_scee))
        s.dihe_scnb = _ptr(f64(topo.dihe_scnb))
        if topo.infr14 is not None and len(topo.infr14):
            s.ninfr14 = len(topo.infr14)
            s.infr14_il = _ptr(i32(np.asarray(topo.infr14).reshape(-1)))
            s.infr14_scee = _ptr(f64(topo.infr14_scee))
            s.infr14_scnb = _ptr(f64(topo.infr14_scnb))
        else:
            s.ninfr14 = 0
        s.igb = topo.igb
        s.dielectric = topo.dielectric
        s.kappa = topo.kappa
        s.gb_offset = 0.09
        s.pb_radii = _ptr(f64(topo.pb_radii))
        s.gb_screen = _ptr(f64(topo.gb_screen))
        a, b, g = topo.gb_alpha_beta_gamma()
        s.gb_alpha = _ptr(f64(a))
        s.gb_beta = _ptr(f64(b))
        s.gb_gamma = _ptr(f64(g))
        s.nrst = 0
        s.rst = None
        if restraints is not None and len(restraints):
            src = np.asarray(restraints).reshape(-1)
            rst = np.zeros(len(src), dtype=ORC_RESTRAINT)
            for name in ORC_RESTRAINT.names:
                rst[name] = src[name]
            keep.append(rst)
            s.nrst = len(rst)
            s.rst = _ptr(rst)
        self.struct = s
        self.keep = keep
        self.natom = n


def eval_mm(topo, xyz: np.ndarray, eval_force: bool = True, nrg_bits: int = 26,
            handle: Optional[SystemHandle] = None, clash_distance: float = 0.0,
            clash_ratio: float = 0.0, restraints=None,
            step: int = 0) -> Tuple[Dict[str, float], Dict[str, int], np.ndarray]:
    """One FP64 evaluation; returns (double energies, fixed-point energies, forces (N, 3)).
    A positive clash_distance or clash_ratio softens Coulomb, Lennard-Jones and 1:4 pairs;
    restraints (or those of `handle`) are mixed at step number `step`."""
    h = handle or SystemHandle(topo, restraints)
    xyz = np.asarray(xyz, dtype=np.float64)
    x = np.ascontiguousarray(xyz[:, 0])
    y = np.ascontiguousarray(xyz[:, 1])
    z = np.ascontiguousarray(xyz[:, 2])
    n = topo.natom
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    e = OrcEnergy()
    lib().orc_eval_mm_ext(ctypes.byref(h.struct), _ptr(x), _ptr(y), _ptr(z), _ptr(fx),
                          _ptr(fy), _ptr(fz), int(eval_force), nrg_bits, float(clash_distance),
                          float(clash_ratio), int(step), ctypes.byref(e))
    nrg = {name: e.nrg[k] for k, name in enumerate(SLOT_NAMES)}
    acc = {name: e.acc[k] for k, name in enumerate(SLOT_NAMES)}
    return nrg, acc, np.stack([fx, fy, fz], axis=1)


TERM = {"BOND": 1, "ANGLE": 2, "DIHEDRAL": 4, "QQ14": 8, "LJ14": 16, "QQ": 32, "LJ": 64,
        "GB": 128, "ALL": 255}


def eval_terms(topo, xyz: np.ndarray, force_terms: Sequence[str], nrg_bits: int = 26,
               handle: Optional[SystemHandle] = None):
    """All energies, with forces from the named terms only (keys of TERM)."""
    h = handle or SystemHandle(topo)
    xyz = np.asarray(xyz, dtype=np.float64)
    x = np.ascontiguousarray(xyz[:, 0])
    y = np.ascontiguousarray(xyz[:, 1])
    z = np.ascontiguousarray(xyz[:, 2])
    n = topo.natom
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    mask = 0
    for t in force_terms:
        mask |= TERM[t]
    e = OrcEnergy()
    lib().orc_eval_terms(ctypes.byref(h.struct), _ptr(x), _ptr(y), _ptr(z), _ptr(fx), _ptr(fy),
                         _ptr(fz), mask, nrg_bits, ctypes.byref(e))
    nrg = {name: e.nrg[k] for k, name in enumerate(SLOT_NAMES)}
    acc = {name: e.acc[k] for k, name in enumerate(SLOT_NAMES)}
    return nrg, acc, np.stack([fx, fy, fz], axis=1)


def eval_gb(topo, xyz: np.ndarray, eval_force: bool = True, nrg_bits: int = 26,
            handle: Optional[SystemHandle] = None):
    """GB only; returns (energy, fixed-point energy, forces, psi, radii, sumdeijda)."""
    h = handle or SystemHandle(topo)
    xyz = np.asarray(xyz, dtype=np.float64)
    x = np.ascontiguousarray(xyz[:, 0])
    y = np.ascontiguousarray(xyz[:, 1])
    z = np.ascontiguousarray(xyz[:, 2])
    n = topo.natom
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    psi, radii, sdi = np.zeros(n), np.zeros(n), np.zeros(n)
    e = OrcEnergy()
    lib().orc_gb(ctypes.byref(h.struct), _ptr(x), _ptr(y), _ptr(z), _ptr(fx), _ptr(fy),
                 _ptr(fz), int(eval_force), float(2.0 ** nrg_bits), ctypes.byref(e), _ptr(psi),
                 _ptr(radii), _ptr(sdi))
    return (e.nrg[SLOT["GENERALIZED_BORN"]], e.acc[SLOT["GENERALIZED_BORN"]],
            np.stack([fx, fy, fz], axis=1), psi, radii, sdi)


def total_potential(nrg: Dict[str, float]) -> float:
    return float(sum(nrg[k] for k in SLOT_NAMES))


def cubic_move(evec: Sequence[float], mvec: Sequence[float]):
    ev = np.ascontiguousarray(evec, dtype=np.float64)
    mv = np.ascontiguousarray(mvec, dtype=np.float64)
    coefs = np.zeros(4)
    nbm = ctypes.c_double(0.0)
    mf = lib().orc_cubic_move(_ptr(ev), _ptr(mv), ctypes.byref(nbm), _ptr(coefs))
    return mf, nbm.value, coefs


# Columns of the per-cycle line-search record of minimize(..., record=True).
LINE_RECORD = ["nrg_a", "nrg_b", "nrg_c", "nrg_d", "mfac_b", "mfac_c", "move_factor", "l_move",
               "gamma"]

# Deliberate mistakes a test may plant in the minimizer (plant_mistake), stormm_oracle.h.
MISTAKES = {"fletcher_reeves": 1, "no_conjugation": 2, "gamma_sign": 3, "first_growth": 4,
            "first_shrink": 5, "second_growth": 6, "normalization_rounding": 7}


def plant_mistake(name: Optional[str]) -> None:
    """Make the oracle's minimizer commit one deliberate mistake (None: none).  Tests only: a
    check that still passes with the mistake planted does not pin what it claims to pin."""
    lib().orc_plant_mistake(0 if name is None else MISTAKES[name])


def minimize(topo, xyz_fixed: np.ndarray, gpos_bits: int, frc_bits: int, nrg_bits: int,
             nsteps: int, sd_cycles: int = 50, dx0: float = 0.01, dp_mode: bool = False,
             handle: Optional[SystemHandle] = None, clash_cycles: int = 0,
             clash_r0: float = 0.2, clash_ratio: float = 0.5, restraints=None,
             record: bool = False):
    """GPU-semantics minimization on fixed-point coordinates; returns (coords, energy trace)
    and, with record=True, also the per-cycle line-search record ([nsteps, 9], columns
    LINE_RECORD).  The first clash_cycles cycles run with clash damping shrinking from
    (clash_r0, clash_ratio) to zero; restraints follow the minimizer's step counters."""
    h = handle or SystemHandle(topo, restraints)
    c = np.array(xyz_fixed, dtype=np.int64).reshape(-1, 3)
    x = np.ascontiguousarray(c[:, 0])
    y = np.ascontiguousarray(c[:, 1])
    z = np.ascontiguousarray(c[:, 2])
    trace = np.zeros(nsteps + 1)
    rec = np.zeros((max(nsteps, 1), len(LINE_RECORD)))
    lib().orc_minimize_record(ctypes.byref(h.struct), _ptr(x), _ptr(y), _ptr(z), gpos_bits,
                              frc_bits, nrg_bits, nsteps, sd_cycles, dx0, int(clash_cycles),
                              float(clash_r0), float(clash_ratio), int(dp_mode), _ptr(trace),
                              _ptr(rec))
    out = np.stack([x, y, z], axis=1)
    if record:
        return out, trace, rec[:nsteps]
    return out, trace


class Batch:
    """Many systems for the threaded CPU baseline."""

    def __init__(self, topos, coords: List[np.ndarray]):
        self.handles = [SystemHandle(t) for t in topos]
        arr_t = OrcSystem * len(topos)
        self.structs = arr_t(*[h.struct for h in self.handles])
        counts = np.array([t.natom for t in topos], dtype=np.int64)
        self.offsets = np.zeros(len(topos), dtype=np.int64)
        self.offsets[1:] = np.cumsum(counts)[:-1]
        allc = np.concatenate([np.asarray(c, dtype=np.float64) for c in coords], axis=0)
        self.x = np.ascontiguousarray(allc[:, 0])
        self.y = np.ascontiguousarray(allc[:, 1])
        self.z = np.ascontiguousarray(allc[:, 2])
        tot = int(counts.sum())
        self.fx, self.fy, self.fz = np.zeros(tot), np.zeros(tot), np.zeros(tot)
        self.nsys = len(topos)

    def run(self, repeats: int = 1, nthreads: int = 1, eval_force: bool = True,
            nrg_bits: int = 26) -> Tuple[float, float]:
        etot = ctypes.c_double(0.0)
        secs = lib().orc_eval_batch(self.nsys, self.structs, _ptr(self.offsets), _ptr(self.x),
                                    _ptr(self.y), _ptr(self.z), _ptr(self.fx), _ptr(self.fy),
                                    _ptr(self.fz), int(eval_force), nrg_bits, repeats, nthreads,
                                    ctypes.byref(etot))
        return secs, etot.value
