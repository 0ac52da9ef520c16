// Host-callable launchers of the device kernels (implemented in the .hip translation units).
#pragma once

#include "internal.h"

namespace stormm_amd {

hipError_t launch_mm_eval(int precision, const MMKit &k, int terms, int eval_force,
                          int eval_energy, hipStream_t stream);
hipError_t launch_fill(long long *a, long long *b, long long *c, size_t n, hipStream_t stream);
// Zeroes three int arrays of n entries (overflow words).
hipError_t launch_fill_int(int *a, int *b, int *c, size_t n, hipStream_t stream);
// ScoreCard::commit: running sums and sums of squares, and (series non-null) the time-series
// row of this sample, series[s][32] for each system s.
hipError_t launch_scorecard_commit(const long long *inst, double *running, double *squared,
                                   long long *series, int nsys, double inv_scale,
                                   hipStream_t stream);

// Energy snapshots of a minimization, taken by the conjugate-gradient kernel from the ScoreCard
// rows of the evaluation that precedes it, every `every` cycles (ScoreCard::commit at the
// reference's diagnostic print frequency).  The snapshot of system s at cycle c goes to sample
// first + c / every, at rows + sample * sample_stride + s * sys_stride, and is added to the
// running sums (when present).  rows == null: no snapshots.
struct HistKit {
  long long *rows;
  double *running, *squared;
  const long long *scorecard;   // instantaneous rows, 32 slots per system
  long long sample_stride, sys_stride;
  int every, first, capacity;   // capacity: samples the rows hold
  int max_cycle;                // no snapshot at or past this cycle (0: no limit)
  double inv_nrg_scale;
};

// Minimizer pieces (mm_minimize.hip).
struct CgKit {
  int nsys;
  const int *slots;       // system of each block (null: block index)
  const int *sys_start, *sys_natom;
  int *steps;             // per-system completed cycles, advanced here (null: use `step`)
  int *done;              // per-system stop flag (null: none)
  double *rms;            // per-system RMS force of the last direction update (optional)
  double conv_rms;        // stop a system when its RMS force falls to this (0 = never)
  int max_steps;          // stop a system after this many cycles (0 = no limit)
  long long *xfrc, *yfrc, *zfrc;
  long long *xprv, *yprv, *zprv;
  long long *xcg, *ycg, *zcg;
  // Overflow words of 95-bit forces, previous forces and directions (frc_bits > 40, double
  // precision only); null otherwise.
  int *xfrc_ovrf, *yfrc_ovrf, *zfrc_ovrf;
  int *xprv_ovrf, *yprv_ovrf, *zprv_ovrf;
  int *xcg_ovrf, *ycg_ovrf, *zcg_ovrf;
  double inv_frc_scale;
  int dp_mode;
  HistKit hist;
};
struct LineKit {
  int nsys;
  const int *slots;       // system of each block (null: block index)
  const int *done;        // per-system stop flag (null: none)
  const int *sys_start, *sys_natom;
  long long *xcrd, *ycrd, *zcrd;
  const long long *xfrc, *yfrc, *zfrc;
  int *xcrd_ovrf, *ycrd_ovrf, *zcrd_ovrf;          // 95-bit coordinates (null: 64-bit)
  const int *xfrc_ovrf, *yfrc_ovrf, *zfrc_ovrf;    // 95-bit directions (null: 64-bit)
  const long long *scorecard;
  double inv_nrg_scale;
  double inv_frc_scale, gpos_scale;
  double *l_move, *s_move, *mfac_a, *mfac_b, *mfac_c, *nrg_a, *nrg_b, *nrg_c, *nrg_d;
  int dp_mode;
};
hipError_t launch_conjugate_gradient(const CgKit &k, int step, int sd_cycles,
                                     hipStream_t stream);
// After a minimization of `cycles` cycles with per-system cycle counts `steps`: the samples of
// a system that stopped early repeat its last snapshot up to sample first + count - 1.
hipError_t launch_history_fill(const HistKit &h, const int *steps, int nsys, int count,
                               hipStream_t stream);
hipError_t launch_line_advance(const LineKit &k, int move_number, hipStream_t stream);

// Fused minimization cycle (mm_minimize.hip): evaluations that make the pending line move of
// the systems they hold while loading the coordinates.  cg.done is required; line holds the
// line-search state and scales.
struct CycleKit {
  CgKit cg;
  LineKit line;
};
// Force + energy (forces) or energy-only evaluation, making line move pending_move (0..3) first
// (-1: none).
hipError_t launch_min_cycle(int precision, const MMKit &k, const CycleKit &ck, bool forces,
                            int pending_move, hipStream_t stream);

// Static per-device state of the allocation-free minimization (mm_minimize.hip): up to
// kNoAllocMaxSystems systems, 16 bytes each (cycle count, stop flag, RMS force).
constexpr int kNoAllocMaxSystems = 65536;
constexpr size_t kNoAllocScratchBytes = static_cast<size_t>(kNoAllocMaxSystems) * 16;
// Device address of the calling thread's current device's copy of that state.
hipError_t noalloc_scratch(void **ptr);

// Fresh per-system state of newly admitted pool slots: layout, zero cycle count and stop flag,
// line-search state primed with dx0, zeroed ScoreCard row.
struct SlotInitKit {
  int count, cap_slots;
  const int *slots, *starts, *natoms;   // per new entry
  int *slot_start, *slot_natom, *steps, *done;
  double *rms, *line;                   // line: 9 arrays of cap_slots
  long long *scorecard;
  double dx0;
};
hipError_t launch_pool_init_slots(const SlotInitKit &k, hipStream_t stream);

}  // namespace stormm_amd
