// Internal device-side layouts shared by the host builder (synthesis.cpp) and the kernels
// (mm_kernels.hip).  Not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/stormm_amd.h"

namespace stormm_amd {

// Tile edge of the all-pairs work: one 16-lane DPP row of a wave64 owns a 16 x 16 tile of
// atom pairs and rotates the j-atom data through its lanes.
constexpr int kTile = 16;
// Systems are laid out in the global per-atom arrays starting on 32-atom boundaries, the
// reference's warp-size padding (src/Synthesis/phasespace_synthesis.cpp, atom_starts).
constexpr int kGlobalPad = 32;
// Largest number of atoms a work unit stages in LDS.
constexpr int kMaxWuAtoms = 512;
// Largest number of systems packed into one work unit.
constexpr int kMaxWuSystems = 32;
// Largest number of tile tasks of one work unit: its systems take at most kMaxWuAtoms / kTile
// atom blocks between them, and b blocks hold b (b + 1) / 2 tiles.
constexpr int kMaxWuTasks = (kMaxWuAtoms / kTile) * ((kMaxWuAtoms / kTile) + 1) / 2;
// Energy components tracked per system inside a work unit (StateVariable slots 0..12).
constexpr int kNrgSlots = 13;
// ScoreCard stride (slots per system), src/Potential/scorecard.cpp:66 with 32-wide warps.
constexpr int kScoreStride = 32;
// StateVariable slots of the kinetic energy and of the potential and total energy sums
// (src/Potential/energy_enumerators.h:36-78).
constexpr int kKineticSlot = 14;
constexpr int kPotentialSlot = 28;
constexpr int kTotalSlot = 29;
// Fixed-precision scales of the LDS accumulators for the Born-radius integral and dE/dR.
constexpr int kPsiBits = 40;
constexpr int kDedrBits = 32;
// LDS coordinate resolution (bits after the binary point) in single-precision mode.
constexpr int kLdsCrdBits = 24;
// Fixed-precision bits of the per-system origins kept in LDS for 95-bit coordinates (values at
// 2^-38 A fit a 64-bit word to 3.3e7 A), and the largest coordinate and force bits that fit a
// plain 64-bit word (src/Constants/fixed_precision.h:37,85).
constexpr int kOriginBits = 38;
constexpr int kMaxPlainGposBits = 38;
constexpr int kMaxPlainFrcBits = 40;
// Threads of an evaluation block: the default, and the largest a single-precision launch may
// use; double-precision launches use at most kMaxEvalBlockDouble threads (their kernels need
// more than the 128 registers per lane that 16-wavefront blocks allow).
constexpr int kEvalBlock = 512;
constexpr int kMaxEvalBlock = 1024;
constexpr int kMaxEvalBlockDouble = 512;
// Entries of each wavefront's LDS queue of near-field Born-integral pairs.  A wave drains 64
// entries at a time whenever the queue holds at least 64 after a group of four rotation steps,
// each of which adds at most 64 (and stores its lanes without an entry just past the new
// ones).
constexpr int kNearQueue = 320;
constexpr int kNearQueueStride = kNearQueue;
constexpr int kStageCounters = 3;  // work-item counters of the psi, derivative and valence stages

// StateVariable slots, src/Potential/energy_enumerators.h:36-75.
enum NrgSlot : int {
  kBond = 0, kAngle = 1, kProperDihedral = 2, kImproperDihedral = 3, kVdw = 7, kVdw14 = 8,
  kElec = 9, kElec14 = 10, kGb = 11, kRestraint = 12
};

// Term selection for the fused kernel.
enum TermMask : int { kTermNonbonded = 1, kTermValence = 2 };

// Valence term kinds (VTerm.x).
enum VKind : int { kVBond = 0, kVAngle = 1, kVDihedral = 2 };

// Work-unit descriptor, 8 ints.  The unit's tile tasks are the clean off-diagonal tasks of all of
// its systems, then their other off-diagonal tasks, then their diagonal tasks (so that most
// wavefronts hold four tasks of one kind: clean tiles run steps without pair masks, and
// wavefronts given only diagonal tiles run the short rotation); its valence terms are those of
// its systems in order.
struct WuDesc {
  int sys_first;     // index into the work-unit system list
  int sys_count;
  int offd_count;    // off-diagonal tile tasks of all systems
  int task_count;    // all tile tasks
  int lds_atoms;     // atoms staged in LDS (each system padded to kTile)
  int term_count;    // valence terms of all systems
  int lj_table;      // 1 = systems in this unit read their stored LJ tables
  int rst_count;     // restraints of all systems
};

// One system inside a work unit and the ranges of its topology template: the template's tile
// tasks (off-diagonal first, then diagonal) and valence terms carry atom indices relative to the
// system's first atom, so every system with the same topology shares one template.
struct WuSys {
  int sys_index;     // system number: ScoreCard row and per-system minimizer state
  int lds_base;      // first LDS atom of the system in the unit
  int global_start;  // first atom in the global per-atom arrays
  int natom;
  int task_first;    // template tile tasks: clean off-diagonal, other off-diagonal, diagonal
  int offd_count;    // off-diagonal tasks, the first clean_count of them clean (16 x 16 atoms,
  int clean_count;   // no excluded pair)
  int diag_count;
  int term_first;    // template valence terms: bonds, then angles, then dihedrals
  int term_count;
  int bond_count;
  int angl_count;
  int lj_off;        // stored LJ table (offset, type count)
  int lj_nt;
  int rst_first;     // the system's restraints in the synthesis restraint list
  int rst_count;
};

// One 16 x 16 tile task of a template: x = first atom of the i tile | (first atom of the j tile
// << 16), relative to the system; y = (diagonal << 8) | (ni << 12) | (nj << 20); z =
// exclusion-mask offset (-1 when the tile holds no excluded pair); w = unused.  When a work unit
// fetches a task it adds the system's LDS base to x and its slot in the unit to y bits 0..7.
using Task = int4;

// Everything the fused kernel reads, passed by value.
// Cosine dihedral parameters: amplitude, cosine and sine of the phase, integer periodicity and
// the improper flag (which energy slot the term reports to).
struct DiheParam {
  double amp, cos_phase, sin_phase;
  int period, improper;
};

// Largest dihedral periodicity evaluated by the multiple-angle expansion.
constexpr int kMaxDihePeriod = 32;

struct MMKit {
  int nsys, nwu, maxa, block;
  int gpos_bits, frc_bits, nrg_bits;
  int igb;
  double dielectric, kappa;
  const int *sys_start, *sys_natom;
  const WuDesc *wu;
  const WuSys *wu_sys;
  const Task *tasks;
  const unsigned *masks;
  const int4 *vterms;
  const double2 *bond_par;   // k, r0
  const double2 *angl_par;   // k, theta0
  const DiheParam *dihe_par;
  const double2 *dihe14;     // 1 / scee (0 = none), 1 / scnb
  const double *q;           // charge * sqrt(Coulomb constant)
  const double *ljp1;        // sigma (combining-rule systems)
  const double *ljp2;        // 2 sqrt(epsilon) (combining-rule systems)
  const int *lj_type;        // type index (table systems)
  const double2 *ljab;       // A, B per system table
  const double2 *ljab14;     // 1:4 A, B per system table
  const double *rho;         // intrinsic radius - offset
  const double *scr;         // screening factor * rho
  const double *pbr;         // intrinsic radius
  const double *gb_alpha, *gb_beta, *gb_gamma;
  long long *xcrd, *ycrd, *zcrd;
  long long *xfrc, *yfrc, *zfrc;
  // Overflow words of 95-bit coordinates (gpos_bits > 38) and forces (stored frc_bits > 40),
  // value = ovrf 2^63 + word; null when the values are plain 64-bit integers.  Forces are
  // accumulated at frc_bits and stored shifted left by frc_store_shift bits.
  const int *xcrd_ovrf, *ycrd_ovrf, *zcrd_ovrf;
  int *xfrc_ovrf, *yfrc_ovrf, *zfrc_ovrf;
  int frc_store_shift;
  long long *scorecard;
  int *status;
  int store_overwrite;       // 1: forces and energy rows are written, not added to
  // Clash damping (src/MolecularMechanics/hpc_minimization.cu:269-281).  clash_variant selects
  // the kernels that can soften Coulomb, Lennard-Jones and 1:4 pairs.  With clash_cycles == 0
  // every system takes clash_r0 / clash_ratio as given; otherwise a system at cycle i takes them
  // scaled by (clash_cycles - i) / clash_cycles while i < clash_cycles, where i is
  // clash_step[system] + clash_step_offset, or clash_cycle when clash_step is null.
  int clash_variant;
  int clash_cycles, clash_cycle, clash_step_offset;
  const int *clash_step;
  double clash_r0, clash_ratio;
  // Restraints (src/Potential/valence_potential.tpp:2090-2160) and the step number their
  // parameters are mixed at: rst_step, or, when rst_step_sys holds each system's cycle count,
  // minimizer_step(rst_step_sys[system] + rst_cycle_offset, rst_move, rst_clash_cycles).
  const stormm_restraint *rst;
  int rst_step;
  const int *rst_step_sys;
  int rst_cycle_offset, rst_move, rst_clash_cycles;
};

// The step counter a restraint sees in the reference's minimizer
// (src/MolecularMechanics/hpc_minimization.cu:262-405).  Each of its four evaluation controls
// counts its own launches: the force evaluation of cycle i runs at step i; the move-th energy
// evaluation (move 1..3) runs at 3 j + move - 1, where j counts the earlier cycles of the same
// kind (clash-damped cycles below clash_cycles, plain cycles after); the final energy evaluation
// after `total` cycles is the plain controls' next launch, minimizer_step(total, 1, clash_cycles).
__host__ __device__ inline int minimizer_step(int cycle, int move, int clash_cycles) {
  if (move == 0) {
    return cycle;
  }
  const int j = (cycle < clash_cycles) ? cycle : cycle - clash_cycles;
  return (3 * j) + move - 1;
}

// Status bits reported by kernels.
enum StatusBit : int { kStatusCoordinateRange = 1, kStatusNonFinite = 2 };

}  // namespace stormm_amd
