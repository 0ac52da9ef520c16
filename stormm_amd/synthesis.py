"""Host-side mirror of the reference's synthesis objects and launch functions over the C-ABI.

Names, argument meaning and error behaviour follow the reference (paths relative to the
reference repository root):

  AtomGraphSynthesis      src/Synthesis/atomgraph_synthesis.h
  PhaseSpaceSynthesis     src/Synthesis/phasespace_synthesis.h
  ScoreCard               src/Potential/scorecard.h
  MolecularMechanicsControls  src/MolecularMechanics/mm_controls.h
  LineMinimization        src/MolecularMechanics/line_minimization.h
  launch_nonbonded        launchNonbonded, src/Potential/hpc_nonbonded_potential.h:142-157
  launch_valence          launchValence, src/Potential/hpc_valence_potential.h:356-378
  launch_conjugate_gradient  launchConjugateGradient, src/Math/hpc_reduction.h:55-63
  launch_line_advance     launchLineAdvance, src/MolecularMechanics/hpc_minimization.h:63-70
  launch_minimization     launchMinimization, src/MolecularMechanics/hpc_minimization.h:122-151

Errors raise StormmError (a RuntimeError), as the reference's rtErr raises std::runtime_error.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from . import restraints as R
from ._lib import (LaunchShape, MinimizeControls, PREC_DOUBLE, PREC_SINGLE, StormmError,
                   SynthesisInfo, TopologyBatch, check)
from .topology import Topology

SCORECARD_STRIDE = 32
STATE_VARIABLES = ["BOND", "ANGLE", "PROPER_DIHEDRAL", "IMPROPER_DIHEDRAL", "UREY_BRADLEY",
                   "CHARMM_IMPROPER", "CMAP", "VDW", "VDW_ONE_FOUR", "ELECTROSTATIC",
                   "ELEC_ONE_FOUR", "GENERALIZED_BORN", "RESTRAINT"]
# StateVariable slots of the potential and total energy sums (src/Potential/
# energy_enumerators.h:36-78).
STATE_POTENTIAL_ENERGY = 28
STATE_TOTAL_ENERGY = 29


def precision_code(prec) -> int:
    if prec in (PREC_SINGLE, "single", "float", "fp32", "f32"):
        return PREC_SINGLE
    if prec in (PREC_DOUBLE, "double", "fp64", "f64"):
        return PREC_DOUBLE
    raise ValueError(f"unknown precision model {prec!r}")


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    return None if a is None else a.ctypes.data


def _stream_ptr(stream) -> Optional[int]:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    handle = getattr(stream, "cuda_stream", None)
    if handle is not None:
        return int(handle)
    return int(stream)


class _BatchArrays:
    """Concatenated per-system topology arrays behind one TopologyBatch struct."""

    def __init__(self, topos: Sequence[Topology], restraints=None):
        if len(topos) == 0:
            raise ValueError("a synthesis needs at least one topology")
        igbs = {t.igb for t in topos}
        if len(igbs) != 1:
            raise ValueError("all topologies in a synthesis must share one implicit-solvent "
                             "model")
        keep: List[np.ndarray] = []

        def cat(arrs, dtype):
            parts = [np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1)) for a in arrs]
            out = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(0, dtype))
            if out.size == 0:
                out = np.zeros(1, dtype=dtype)
            keep.append(out)
            return out

        def counts(vals):
            out = np.ascontiguousarray(np.asarray(vals, dtype=np.int32))
            keep.append(out)
            return out

        b = TopologyBatch()
        b.system_count = len(topos)
        b.atom_counts = _ptr(counts([t.natom for t in topos]))
        b.charge = _ptr(cat([t.charge for t in topos], np.float64))
        b.lj_type = _ptr(cat([t.lj_idx for t in topos], np.int32))
        b.lj_type_counts = _ptr(counts([t.n_lj_types for t in topos]))
        # Systems with per-type Lorentz-Berthelot parameters send those; the others send their
        # T x T tables.
        comb = [t.lj_sigma is not None for t in topos]
        tabled = [t.lj_tables() for t, c in zip(topos, comb) if not c]
        b.lj_combining = _ptr(counts([1 if c else 0 for c in comb]))
        b.lj_sigma = _ptr(cat([t.lj_sigma for t, c in zip(topos, comb) if c], np.float64))
        b.lj_epsilon = _ptr(cat([t.lj_epsilon for t, c in zip(topos, comb) if c], np.float64))
        b.lja = _ptr(cat([x[0] for x in tabled], np.float64))
        b.ljb = _ptr(cat([x[1] for x in tabled], np.float64))
        b.lja14 = _ptr(cat([x[2] for x in tabled], np.float64))
        b.ljb14 = _ptr(cat([x[3] for x in tabled], np.float64))
        b.pb_radii = _ptr(cat([t.pb_radii for t in topos], np.float64))
        b.gb_screen = _ptr(cat([t.gb_screen for t in topos], np.float64))
        b.bond_counts = _ptr(counts([len(t.bonds) for t in topos]))
        b.bond_atoms = _ptr(cat([t.bonds for t in topos], np.int32))
        b.bond_k = _ptr(cat([t.bond_k for t in topos], np.float64))
        b.bond_r0 = _ptr(cat([t.bond_r0 for t in topos], np.float64))
        b.angle_counts = _ptr(counts([len(t.angles) for t in topos]))
        b.angle_atoms = _ptr(cat([t.angles for t in topos], np.int32))
        b.angle_k = _ptr(cat([t.angle_k for t in topos], np.float64))
        b.angle_theta = _ptr(cat([t.angle_theta for t in topos], np.float64))
        b.dihedral_counts = _ptr(counts([len(t.dihedrals) for t in topos]))
        b.dihedral_atoms = _ptr(cat([t.dihedrals for t in topos], np.int32))
        b.dihe_amp = _ptr(cat([t.dihe_amp for t in topos], np.float64))
        b.dihe_phase = _ptr(cat([t.dihe_phase for t in topos], np.float64))
        b.dihe_freq = _ptr(cat([t.dihe_freq for t in topos], np.float64))
        b.dihe_improper = _ptr(cat([t.dihe_improper for t in topos], np.int32))
        b.dihe_scee = _ptr(cat([t.dihe_scee for t in topos], np.float64))
        b.dihe_scnb = _ptr(cat([t.dihe_scnb for t in topos], np.float64))
        # Exclusions are the pairs within three bonds, which the planner derives from the bonds
        # once per distinct topology (null exclusion arrays).
        b.exclusion_counts = None
        b.exclusion_pairs = None
        b.igb = topos[0].igb
        b.dielectric = topos[0].dielectric
        b.salt_concentration = topos[0].salt
        b.coulomb_constant = topos[0].coulomb
        # Per-system restraint records (None: no restraints anywhere).
        b.restraint_counts = None
        b.restraints = None
        if restraints is not None:
            if len(restraints) != len(topos):
                raise ValueError("give one restraint array (or None) per system")
            recs = [R.empty() if r is None else np.asarray(r, dtype=R.RESTRAINT_DTYPE).reshape(-1)
                    for r in restraints]
            allr = np.ascontiguousarray(np.concatenate(recs) if recs else R.empty())
            if allr.size == 0:
                allr = R.empty(1)
            keep.append(allr)
            b.restraint_counts = _ptr(counts([len(r) for r in recs]))
            b.restraints = allr.ctypes.data
        self.struct = b
        self.keep = keep


def _shape(shape) -> Optional[LaunchShape]:
    if shape is None:
        return None
    if isinstance(shape, LaunchShape):
        return shape
    s = LaunchShape()
    s.wu_atom_capacity = int(shape.get("wu_atom_capacity", 0))
    s.target_wu_per_cu = int(shape.get("target_wu_per_cu", 0))
    s.block_threads = int(shape.get("block_threads", 0))
    return s


def plan_synthesis(topos: Sequence[Topology], shape=None, restraints=None) -> SynthesisInfo:
    """Host-only work-unit plan (no device memory): counts of units, tasks, pairs, terms."""
    L = _lib.load()
    arrs = _BatchArrays(topos, restraints)
    info = SynthesisInfo()
    sh = _shape(shape)
    check(L.stormm_synthesis_plan(ctypes.byref(arrs.struct),
                                  ctypes.byref(sh) if sh is not None else None,
                                  ctypes.byref(info)))
    return info


def autotune(topos: Sequence[Topology], precision="single", coords=None):
    """Time candidate work-unit shapes on the device (on `coords` when given); returns
    (shape dict, mean ms)."""
    L = _lib.load()
    arrs = _BatchArrays(topos)
    best = LaunchShape()
    ms = ctypes.c_double(0.0)
    xyz = None
    if coords is not None:
        xyz = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1, 3)
                                                   for c in coords]))
    check(L.stormm_autotune_launch_shape(ctypes.byref(arrs.struct), precision_code(precision),
                                         None if xyz is None else xyz.ctypes.data,
                                         ctypes.byref(best), ctypes.byref(ms)))
    return ({"wu_atom_capacity": best.wu_atom_capacity,
             "target_wu_per_cu": best.target_wu_per_cu,
             "block_threads": best.block_threads}, ms.value)


class AtomGraphSynthesis:
    """Many topologies batched into one device-resident parameter set with work units.
    `restraints`, when given, holds one restraint record array (stormm_amd.restraints) or None
    per system, the counterpart of the RestraintApparatus list the reference's synthesis takes
    (src/Synthesis/atomgraph_synthesis.h:58-104)."""

    def __init__(self, topos: Sequence[Topology], shape=None, restraints=None):
        self._L = _lib.load()
        self.topologies = list(topos)
        arrs = _BatchArrays(self.topologies, restraints)
        sh = _shape(shape)
        h = ctypes.c_void_p()
        check(self._L.stormm_synthesis_create(ctypes.byref(arrs.struct),
                                              ctypes.byref(sh) if sh is not None else None,
                                              ctypes.byref(h)))
        self.handle = h
        info = SynthesisInfo()
        check(self._L.stormm_synthesis_get_info(h, ctypes.byref(info)))
        self.info = info
        n = info.system_count
        self.atom_starts = np.zeros(n, dtype=np.int32)
        self.atom_counts = np.zeros(n, dtype=np.int32)
        check(self._L.stormm_synthesis_atom_layout(h, _ptr(self.atom_starts),
                                                   _ptr(self.atom_counts)))

    @property
    def system_count(self) -> int:
        return self.info.system_count

    @property
    def total_atoms(self) -> int:
        return int(self.atom_counts.sum())

    def status(self, clear: bool = True) -> int:
        st = ctypes.c_int32(0)
        check(self._L.stormm_synthesis_status(self.handle, ctypes.byref(st), int(clear)))
        return st.value

    def minimization_status(self):
        cyc = np.zeros(self.system_count, dtype=np.int32)
        rms = np.zeros(self.system_count, dtype=np.float64)
        check(self._L.stormm_minimization_status(self.handle, _ptr(cyc), _ptr(rms)))
        return cyc, rms

    def close(self):
        if getattr(self, "handle", None):
            self._L.stormm_synthesis_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PhaseSpaceSynthesis:
    """Fixed-precision coordinates and forces of every system (int64 at gpos / frc bits)."""

    def __init__(self, ag: AtomGraphSynthesis, gpos_bits: int = 32, frc_bits: int = 23):
        self._L = _lib.load()
        self.ag = ag
        self.gpos_bits = gpos_bits
        self.frc_bits = frc_bits
        h = ctypes.c_void_p()
        check(self._L.stormm_phasespace_create(ag.handle, gpos_bits, frc_bits, ctypes.byref(h)))
        self.handle = h

    def _flat(self, coords) -> np.ndarray:
        if isinstance(coords, (list, tuple)):
            coords = np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1, 3)
                                     for c in coords])
        a = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).reshape(-1, 3))
        if a.shape[0] != self.ag.total_atoms:
            raise ValueError(f"expected {self.ag.total_atoms} atoms, got {a.shape[0]}")
        return a

    def upload(self, coords) -> None:
        a = self._flat(coords)
        check(self._L.stormm_phasespace_upload_coordinates(self.handle, _ptr(a)))

    def upload_fixed(self, fixed: np.ndarray) -> None:
        a = np.ascontiguousarray(np.asarray(fixed, dtype=np.int64).reshape(-1, 3))
        check(self._L.stormm_phasespace_upload_fixed(self.handle, _ptr(a)))

    def _split(self, flat: np.ndarray) -> List[np.ndarray]:
        out, c = [], 0
        for n in self.ag.atom_counts:
            out.append(flat[c:c + n])
            c += n
        return out

    def coordinates(self) -> List[np.ndarray]:
        a = np.zeros((self.ag.total_atoms, 3))
        check(self._L.stormm_phasespace_download_coordinates(self.handle, _ptr(a)))
        return self._split(a)

    def forces(self) -> List[np.ndarray]:
        a = np.zeros((self.ag.total_atoms, 3))
        check(self._L.stormm_phasespace_download_forces(self.handle, _ptr(a)))
        return self._split(a)

    def fixed(self):
        c = np.zeros((self.ag.total_atoms, 3), dtype=np.int64)
        f = np.zeros((self.ag.total_atoms, 3), dtype=np.int64)
        check(self._L.stormm_phasespace_download_fixed(self.handle, _ptr(c), _ptr(f)))
        return self._split(c), self._split(f)

    def fixed_values(self):
        """Coordinates and forces as exact Python integers (object arrays): value =
        overflow word * 2^63 + word, the 95-bit form above 38 coordinate / 40 force bits."""
        n = self.ag.total_atoms
        c, co = np.zeros((n, 3), dtype=np.int64), np.zeros((n, 3), dtype=np.int32)
        f, fo = np.zeros((n, 3), dtype=np.int64), np.zeros((n, 3), dtype=np.int32)
        check(self._L.stormm_phasespace_download_fixed_split(self.handle, _ptr(c), _ptr(co),
                                                             _ptr(f), _ptr(fo)))
        big = lambda w, o: (o.astype(object) * (1 << 63)) + w.astype(object)
        return self._split(big(c, co)), self._split(big(f, fo))

    def writer(self) -> "_lib.PsWriter":
        """The device arrays in the PsSynthesisWriter layout."""
        w = _lib.PsWriter()
        check(self._L.stormm_phasespace_writer(self.handle, ctypes.byref(w)))
        return w

    def _range_atoms(self, lo: int, hi: int) -> int:
        counts = self.ag.atom_counts
        if not (0 <= lo < hi <= len(counts)):
            raise ValueError(f"system range [{lo}, {hi}) is empty or outside the synthesis's "
                             f"{len(counts)} systems")
        return int(sum(counts[lo:hi]))

    def upload_systems(self, lo: int, hi: int, coords) -> None:
        """Coordinates of systems [lo, hi) only (PhaseSpaceSynthesis::upload(kind, lower, upper),
        src/Synthesis/phasespace_synthesis.cpp:2327-2333); the other systems keep theirs."""
        n = self._range_atoms(lo, hi)
        if isinstance(coords, (list, tuple)):
            coords = np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1, 3)
                                     for c in coords])
        a = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).reshape(-1, 3))
        if a.shape[0] != n:
            raise ValueError(f"systems [{lo}, {hi}) hold {n} atoms, got {a.shape[0]}")
        check(self._L.stormm_phasespace_upload_systems(self.handle, lo, hi, _ptr(a)))

    def download_systems(self, lo: int, hi: int):
        """Coordinates and forces of systems [lo, hi) (PhaseSpaceSynthesis::download(kind, lower,
        upper), phasespace_synthesis.cpp:2379-2385), one array per system."""
        n = self._range_atoms(lo, hi)
        c = np.zeros((n, 3))
        f = np.zeros((n, 3))
        check(self._L.stormm_phasespace_download_systems(self.handle, lo, hi, _ptr(c), _ptr(f)))
        cuts = np.cumsum(self.ag.atom_counts[lo:hi])[:-1]
        return np.split(c, cuts), np.split(f, cuts)

    def import_system(self, system_index: int, coords) -> None:
        """PhaseSpaceSynthesis::importSystem (phasespace_synthesis.cpp:3410-3436): one system's
        coordinates, written to the device."""
        self.upload_systems(system_index, system_index + 1, coords)

    def export_system(self, system_index: int):
        """PhaseSpaceSynthesis::exportSystem (phasespace_synthesis.cpp:2472-): one system's
        coordinates and forces, read from the device."""
        c, f = self.download_systems(system_index, system_index + 1)
        return c[0], f[0]

    def initialize_forces(self, stream=None) -> None:
        check(self._L.stormm_phasespace_initialize_forces(self.handle, _stream_ptr(stream)))

    def prime_conjugate_gradient_calculation(self, stream=None) -> None:
        check(self._L.stormm_phasespace_prime_conjugate_gradient(self.handle,
                                                                 _stream_ptr(stream)))

    def close(self):
        if getattr(self, "handle", None):
            self._L.stormm_phasespace_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ScoreCard:
    """Per-system int64 energy accumulators, 32 StateVariable slots per system."""

    def __init__(self, system_count: int, nrg_bits: int = 26):
        self._L = _lib.load()
        self.system_count = system_count
        self.nrg_bits = nrg_bits
        h = ctypes.c_void_p()
        check(self._L.stormm_scorecard_create(system_count, nrg_bits, ctypes.byref(h)))
        self.handle = h

    def initialize(self, stream=None) -> None:
        check(self._L.stormm_scorecard_initialize(self.handle, _stream_ptr(stream)))

    def commit(self, stream=None) -> None:
        check(self._L.stormm_scorecard_commit(self.handle, _stream_ptr(stream)))

    def accumulators(self) -> np.ndarray:
        a = np.zeros((self.system_count, SCORECARD_STRIDE), dtype=np.int64)
        check(self._L.stormm_scorecard_download(self.handle, _ptr(a)))
        return a

    def report_instantaneous_states(self) -> np.ndarray:
        """Energies in kcal/mol, (nsys, 13) in StateVariable order BOND..RESTRAINT."""
        return self.accumulators()[:, :13].astype(np.float64) * 2.0 ** -self.nrg_bits

    def potential_energies(self) -> np.ndarray:
        acc = self.accumulators()[:, :13].sum(axis=1)
        return acc.astype(np.float64) * 2.0 ** -self.nrg_bits

    def history(self):
        n = self.system_count * SCORECARD_STRIDE
        run = np.zeros(n)
        sq = np.zeros(n)
        cnt = ctypes.c_int32(0)
        check(self._L.stormm_scorecard_download_history(self.handle, _ptr(run), _ptr(sq),
                                                        ctypes.byref(cnt)))
        return (run.reshape(-1, SCORECARD_STRIDE), sq.reshape(-1, SCORECARD_STRIDE), cnt.value)

    # ---- time series (ScoreCard::time_series_accumulators, src/Potential/scorecard.h:503) ----
    def sample_count(self) -> int:
        """ScoreCard::getSampleSize (scorecard.h:137)."""
        n, cap = ctypes.c_int32(0), ctypes.c_int32(0)
        check(self._L.stormm_scorecard_sample_count(self.handle, ctypes.byref(n),
                                                    ctypes.byref(cap)))
        return n.value

    def sample_capacity(self) -> int:
        """ScoreCard::getSampleCapacity (scorecard.h:145)."""
        n, cap = ctypes.c_int32(0), ctypes.c_int32(0)
        check(self._L.stormm_scorecard_sample_count(self.handle, ctypes.byref(n),
                                                    ctypes.byref(cap)))
        return cap.value

    def reserve(self, capacity: int) -> None:
        check(self._L.stormm_scorecard_reserve(self.handle, int(capacity)))

    def reset_sample_count(self, count: int = 0) -> None:
        check(self._L.stormm_scorecard_reset_sample_count(self.handle, int(count)))

    def set_last_time_step(self, time_step: int) -> None:
        check(self._L.stormm_scorecard_set_last_time_step(self.handle, int(time_step)))

    def series(self, first: int = 0, count: Optional[int] = None):
        """Samples [first, first + count) as ([count][nsys][32] int64 at nrg_bits, time steps);
        slot 28 holds the potential energy sum and slot 29 the total."""
        if count is None:
            count = self.sample_count() - first
        rows = np.zeros((count, self.system_count, SCORECARD_STRIDE), dtype=np.int64)
        steps = np.zeros(count, dtype=np.int32)
        check(self._L.stormm_scorecard_download_series(self.handle, int(first), int(count),
                                                       _ptr(rows), _ptr(steps)))
        return rows, steps

    def report_history(self, system_index: int = 0, aspect: int = STATE_TOTAL_ENERGY):
        """ScoreCard::reportHistory(aspect, system_index) (scorecard.cpp:1090-1102): one value
        per sample, kcal/mol."""
        rows, _ = self.series()
        return rows[:, system_index, aspect].astype(np.float64) * 2.0 ** -self.nrg_bits

    def report_history_over_systems(self, aspect: int = STATE_TOTAL_ENERGY):
        """ScoreCard::reportHistory(aspect) (scorecard.cpp:1110-1134): per sample, the mean and
        the standard deviation (n - 1 denominator) over the systems."""
        rows, _ = self.series()
        v = rows[:, :, aspect].astype(np.float64) * 2.0 ** -self.nrg_bits
        std = v.std(axis=1, ddof=1) if self.system_count > 1 else np.zeros(len(v))
        return v.mean(axis=1), std

    def report_average_states(self):
        """ScoreCard::reportAverageStates (scorecard.cpp:869-896): running sums over the sample
        count, [nsys][32] kcal/mol."""
        run, _, n = self.history()
        return run / n

    def report_variance_of_states(self):
        """ScoreCard::reportVarianceOfStates (scorecard.cpp:988-1021): sqrt(n S2 - S1^2) /
        sqrt(n (n - 1)) of the running sums S1 and S2, [nsys][32]."""
        run, sq, n = self.history()
        return np.sqrt(np.maximum(n * sq - run * run, 0.0)) / np.sqrt(n * (n - 1.0))

    def close(self):
        if getattr(self, "handle", None):
            self._L.stormm_scorecard_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def minimize_controls(total_cycles: int = 200, sd_cycles: int = 50, dx0: float = 0.01,
                      rms_force: float = 0.0, clash_cycles: Optional[int] = None,
                      clash_r0: float = 0.2, clash_vdw_ratio: float = 0.5,
                      ntpr: int = 0) -> MinimizeControls:
    """Defaults of the reference's &minimize namelist (src/Namelists/nml_minimize.h:21-30):
    maxcyc 200, ncyc 50, cdcyc 25 (capped at maxcyc), dx0 0.01, clash_r0 0.2,
    clash_vdw_ratio 0.5, ntpr 0.  ntpr > 0 commits the ScoreCard every ntpr cycles."""
    m = MinimizeControls()
    m.diagnostic_print_frequency = ntpr
    m.total_cycles = total_cycles
    m.steepest_descent_cycles = sd_cycles
    m.initial_step = dx0
    m.convergence_rms_force = rms_force
    m.clash_damping_cycles = min(25, total_cycles) if clash_cycles is None else clash_cycles
    m.clash_minimum_distance = clash_r0
    m.clash_vdw_ratio = clash_vdw_ratio
    return m


class MolecularMechanicsControls:
    def __init__(self, mincon: Optional[MinimizeControls] = None):
        self._L = _lib.load()
        self.mincon = mincon or minimize_controls()
        h = ctypes.c_void_p()
        check(self._L.stormm_mm_controls_create(ctypes.byref(self.mincon), ctypes.byref(h)))
        self.handle = h

    @property
    def step(self) -> int:
        v = ctypes.c_int32(0)
        check(self._L.stormm_mm_controls_get_step(self.handle, ctypes.byref(v)))
        return v.value

    @step.setter
    def step(self, value: int) -> None:
        check(self._L.stormm_mm_controls_set_step(self.handle, int(value)))

    def close(self):
        if getattr(self, "handle", None):
            self._L.stormm_mm_controls_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LineMinimization:
    FIELDS = ["l_move", "s_move", "mfac_a", "mfac_b", "mfac_c", "nrg_a", "nrg_b", "nrg_c",
              "nrg_d"]

    def __init__(self, system_count: int, dx0: float = 0.01):
        self._L = _lib.load()
        self.system_count = system_count
        h = ctypes.c_void_p()
        check(self._L.stormm_line_min_create(system_count, dx0, ctypes.byref(h)))
        self.handle = h

    def values(self) -> dict:
        a = np.zeros(9 * self.system_count)
        check(self._L.stormm_line_min_download(self.handle, _ptr(a)))
        a = a.reshape(9, self.system_count)
        return {k: a[i] for i, k in enumerate(self.FIELDS)}

    def close(self):
        if getattr(self, "handle", None):
            self._L.stormm_line_min_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _h(obj):
    return None if obj is None else obj.handle


def launch_nonbonded(prec, ag, ctrl, ps, sc, eval_force=True, eval_energy=True, stream=None,
                     clash_distance=0.0, clash_ratio=0.0):
    check(_lib.load().stormm_launch_nonbonded(precision_code(prec), ag.handle, _h(ctrl),
                                              ps.handle, _h(sc), int(eval_force),
                                              int(eval_energy), float(clash_distance),
                                              float(clash_ratio), _stream_ptr(stream)))


def launch_valence(prec, ag, ctrl, ps, sc, eval_force=True, eval_energy=True, stream=None,
                   clash_distance=0.0, clash_ratio=0.0):
    check(_lib.load().stormm_launch_valence(precision_code(prec), ag.handle, _h(ctrl),
                                            ps.handle, _h(sc), int(eval_force),
                                            int(eval_energy), float(clash_distance),
                                            float(clash_ratio), _stream_ptr(stream)))


def launch_mm_evaluation(prec, ag, ctrl, ps, sc, eval_force=True, eval_energy=True,
                         stream=None, clash_distance=0.0, clash_ratio=0.0):
    check(_lib.load().stormm_launch_mm_evaluation(precision_code(prec), ag.handle, _h(ctrl),
                                                  ps.handle, _h(sc), int(eval_force),
                                                  int(eval_energy), float(clash_distance),
                                                  float(clash_ratio), _stream_ptr(stream)))


def launch_conjugate_gradient(prec, ag, ps, ctrl, stream=None):
    check(_lib.load().stormm_launch_conjugate_gradient(precision_code(prec), ag.handle,
                                                       ps.handle, ctrl.handle,
                                                       _stream_ptr(stream)))


def launch_line_advance(prec, ag, ps, sc, line_record, move_number, stream=None):
    check(_lib.load().stormm_launch_line_advance(precision_code(prec), ag.handle, ps.handle,
                                                 sc.handle, line_record.handle, int(move_number),
                                                 _stream_ptr(stream)))


def launch_minimization(prec, ag, ps, mincon, sc, line_record, persistent=True, stream=None):
    check(_lib.load().stormm_launch_minimization(precision_code(prec), ag.handle, ps.handle,
                                                 ctypes.byref(mincon), sc.handle,
                                                 line_record.handle, int(persistent),
                                                 _stream_ptr(stream)))


def time_mm_evaluation(prec, ag, ps, sc, eval_force=True, repeats=10, stream=None) -> float:
    ms = ctypes.c_double(0.0)
    check(_lib.load().stormm_time_mm_evaluation(precision_code(prec), ag.handle, ps.handle,
                                                sc.handle, int(eval_force), int(repeats),
                                                _stream_ptr(stream), ctypes.byref(ms)))
    return ms.value


TERMS = {"nonbonded": 1, "valence": 2, "all": 3}


def time_launch(prec, ag, ps, sc, terms="all", eval_force=True, repeats=10, stream=None) -> float:
    """Mean kernel time (ms) of `repeats` launches of the chosen terms: "nonbonded" (Coulomb,
    Lennard-Jones and GB), "valence" or "all" (the fused evaluation)."""
    ms = ctypes.c_double(0.0)
    check(_lib.load().stormm_time_launch(precision_code(prec), ag.handle, ps.handle, sc.handle,
                                         TERMS[terms], int(eval_force), int(repeats),
                                         _stream_ptr(stream), ctypes.byref(ms)))
    return ms.value


def copy_bandwidth(nbytes: int = 1 << 30, repeats: int = 20) -> float:
    """Measured stream-copy bandwidth (read + write) in GB/s."""
    v = ctypes.c_double(0.0)
    check(_lib.load().stormm_copy_bandwidth(int(nbytes), int(repeats), ctypes.byref(v)))
    return v.value


def fma_throughput(iters: int = 65536) -> float:
    """Measured non-packed FP32 FMA rate in TFLOP/s."""
    v = ctypes.c_double(0.0)
    check(_lib.load().stormm_fma_throughput(int(iters), ctypes.byref(v)))
    return v.value


def synchronize() -> None:
    check(_lib.load().stormm_device_synchronize())


def device_info():
    cu = ctypes.c_int32(0)
    lds = ctypes.c_int32(0)
    name = ctypes.create_string_buffer(256)
    check(_lib.load().stormm_device_info(ctypes.byref(cu), ctypes.byref(lds), name, 256))
    return {"cu_count": cu.value, "lds_bytes_per_cu": lds.value, "name": name.value.decode()}
