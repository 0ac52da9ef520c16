// Line-minimization kernels: the conjugate-gradient direction update and the four line moves of
// the reference's launch-per-step minimizer, plus the persistent minimizer that runs every cycle
// of a work unit's systems inside one block.
#include <algorithm>
#include <mutex>

#include "minimize_device.h"
#include "mm_eval.h"
#include "kernels.h"

namespace stormm_amd {

__device__ __forceinline__ double wave_sum64(double v) {
  for (int m = 32; m >= 1; m >>= 1) {
    v += __shfl_xor(v, m, 64);
  }
  return v;
}

// Polak-Ribiere direction and normalization, one 64-lane wave per system:
// src/Math/hpc_reduction.cu:39-244 (gather / scatter cores) and
// src/Math/conjugate_gradient.cui:259-370 (all-reduce kernel).  The per-lane accumulation order
// and the butterfly reductions match the persistent minimizer, so both paths produce the same
// directions bit for bit.
//
// With per-system step counters (serving), each system advances its own counter and stops (its
// `done` flag is set, the direction left untouched and no line move follows) once its RMS force
// reaches `conv_rms` or its counter reaches `max_steps`: the persistent minimizer's stopping rule.
// Systems of up to kCgCached x 64 atoms keep their forces and directions in registers from the
// first pass to the last, so the three passes cost one round trip to memory; larger systems read
// them again in each pass.  Both forms accumulate in the same order.
constexpr int kCgCached = 4;

// Per-system stop rule of the serving and graph minimizers: records the RMS force, and either
// stops the system (returns false) or advances its cycle count into *step.
__device__ __forceinline__ bool cg_continue(const CgKit &k, int s, int lane, int n, double ff,
                                            int *step) {
  if (k.steps == nullptr) {
    return true;
  }
  const int prior = k.steps[s];
  const double rms = sqrt(ff / (3.0 * n)) * k.inv_frc_scale;
  if (lane == 0 && k.rms != nullptr) {
    k.rms[s] = rms;
  }
  const bool stop = (k.conv_rms > 0.0 && rms <= k.conv_rms) ||
                    (k.max_steps > 0 && prior >= k.max_steps);
  if (stop) {
    if (lane == 0) {
      k.done[s] = 1;
    }
    return false;
  }
  *step = prior + 1;
  if (lane == 0) {
    k.steps[s] = *step;
  }
  return true;
}

// The same update on 95-bit forces, previous forces and directions (double precision with more
// than 40 force bits): values y 2^63 + x are read as doubles, new directions rounded to nearest
// and split as doubleToInt95 does (src/Math/hpc_reduction.cu:95-321, the overflow forms).
__device__ void cg_wide(const CgKit &k, int s, int lane, int start, int n, int step,
                        int sd_cycles) {
  long long *frc[3] = {k.xfrc, k.yfrc, k.zfrc};
  long long *prv[3] = {k.xprv, k.yprv, k.zprv};
  long long *cg[3] = {k.xcg, k.ycg, k.zcg};
  int *frc_o[3] = {k.xfrc_ovrf, k.yfrc_ovrf, k.zfrc_ovrf};
  int *prv_o[3] = {k.xprv_ovrf, k.yprv_ovrf, k.zprv_ovrf};
  int *cg_o[3] = {k.xcg_ovrf, k.ycg_ovrf, k.zcg_ovrf};
  double ff = 0.0, gg = 0.0, dgg = 0.0;
  for (int i = lane; i < n; i += 64) {
    const int g = start + i;
    for (int d = 0; d < 3; d++) {
      const __int128 f = i95_value(frc[d][g], frc_o[d][g]);
      const __int128 p = i95_value(prv[d][g], prv_o[d][g]);
      const double df = i128_to_double(f);
      const double dp = i128_to_double(p);
      ff += df * df;
      gg += dp * dp;
      dgg += i128_to_double(f - p) * df;
    }
  }
  ff = wave_sum64(ff);
  gg = wave_sum64(gg);
  dgg = wave_sum64(dgg);
  if (!cg_continue(k, s, lane, n, ff, &step)) {
    return;
  }
  const double gam = (step >= sd_cycles) ? (dgg / gg) : 0.0;
  const bool use_gam = fabs(gam) > 1.0e-12;
  double msum = 0.0;
  for (int i = lane; i < n; i += 64) {
    const int g = start + i;
    for (int d = 0; d < 3; d++) {
      const long long fx = frc[d][g];
      const int fy = frc_o[d][g];
      const double df = i95_to_double(fx, fy);
      prv[d][g] = fx;
      prv_o[d][g] = fy;
      if (use_gam) {
        const double dfcg = df + (gam * i95_to_double(cg[d][g], cg_o[d][g]));
        msum += dfcg * dfcg;
        long long x;
        int y;
        i95_from_double(dfcg, &x, &y);
        cg[d][g] = x;
        cg_o[d][g] = y;
        frc[d][g] = x;
        frc_o[d][g] = y;
      }
      else {
        msum += df * df;
        cg[d][g] = fx;
        cg_o[d][g] = fy;
      }
    }
  }
  msum = wave_sum64(msum);
  const double inv_factor = 1.0 / (sqrt(msum) * k.inv_frc_scale);
  for (int i = lane; i < n; i += 64) {
    const int g = start + i;
    for (int d = 0; d < 3; d++) {
      long long x;
      int y;
      i95_from_double(i95_to_double(frc[d][g], frc_o[d][g]) * inv_factor, &x, &y);
      frc[d][g] = x;
      frc_o[d][g] = y;
    }
  }
}

// Sample `sample` of system s in the energy history: its ScoreCard row with the potential
// energy sum in slot 28 and the total in slot 29, added to the running sums when they are kept
// (ScoreCard::commit, src/Potential/scorecard.cpp:366-396, and its device form, src/Potential/
// hpc_scorecard.cu:170-215).  Lanes 0..31 take one slot each.
__device__ __forceinline__ void write_snapshot(const HistKit &h, int s, int sample,
                                               const long long *row, int lane) {
  if (lane >= kScoreStride) {
    return;
  }
  long long pe = 0;
  for (int k = 0; k < kNrgSlots; k++) {
    pe += row[k];
  }
  const long long v = (lane == kPotentialSlot) ? pe :
                      (lane == kTotalSlot) ? pe + row[kKineticSlot] : row[lane];
  h.rows[static_cast<long long>(sample) * h.sample_stride +
         static_cast<long long>(s) * h.sys_stride + lane] = v;
  if (h.running != nullptr) {
    const double d = static_cast<double>(v) * h.inv_nrg_scale;
    const size_t at = static_cast<size_t>(s) * kScoreStride + lane;
    h.running[at] += d;
    h.squared[at] += d * d;
  }
}

// The snapshot of cycle `cycle` (0-based), taken when the cycle is a multiple of the history's
// interval (src/MolecularMechanics/hpc_minimization.cu:292-300).
__device__ __forceinline__ void record_snapshot(const HistKit &h, int s, int cycle, int lane) {
  if (h.rows == nullptr || (cycle % h.every) != 0 ||
      (h.max_cycle > 0 && cycle >= h.max_cycle)) {
    return;
  }
  const int sample = h.first + (cycle / h.every);
  if (sample < h.capacity) {
    write_snapshot(h, s, sample, h.scorecard + static_cast<size_t>(s) * kScoreStride, lane);
  }
}

__global__ void __launch_bounds__(64) k_conjugate_gradient(CgKit k, int step, int sd_cycles) {
  const int s = (k.slots != nullptr) ? k.slots[blockIdx.x] : static_cast<int>(blockIdx.x);
  const int lane = threadIdx.x;
  if (k.done != nullptr && k.done[s] != 0) {
    return;
  }
  // The evaluation before this kernel left the energies at the start of the cycle in the
  // ScoreCard: cycles counted on the device, or `step` = cycle + 1 from the host.
  record_snapshot(k.hist, s, (k.steps != nullptr) ? k.steps[s] : step - 1, lane);
  const int start = k.sys_start[s];
  const int n = k.sys_natom[s];
  if (k.xfrc_ovrf != nullptr) {
    cg_wide(k, s, lane, start, n, step, sd_cycles);
    return;
  }
  long long *frc[3] = {k.xfrc, k.yfrc, k.zfrc};
  long long *prv[3] = {k.xprv, k.yprv, k.zprv};
  long long *cg[3] = {k.xcg, k.ycg, k.zcg};
  const bool cached = (n <= 64 * kCgCached);
  long long rf[kCgCached][3], rp[kCgCached][3], rc[kCgCached][3];
  if (cached) {
#pragma unroll
    for (int it = 0; it < kCgCached; it++) {
      const int i = lane + (64 * it);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        rf[it][d] = (i < n) ? frc[d][start + i] : 0LL;
        rp[it][d] = (i < n) ? prv[d][start + i] : 0LL;
        rc[it][d] = (i < n) ? cg[d][start + i] : 0LL;
      }
    }
  }
  double ff = 0.0, gg = 0.0, dgg = 0.0;
  auto accumulate = [&](long long lf, long long lp) {
    const double df = static_cast<double>(lf);
    const double dp = static_cast<double>(lp);
    ff += df * df;
    gg += dp * dp;
    dgg += static_cast<double>(lf - lp) * df;
  };
  if (cached) {
#pragma unroll
    for (int it = 0; it < kCgCached; it++) {
      if (lane + (64 * it) < n) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
          accumulate(rf[it][d], rp[it][d]);
        }
      }
    }
  }
  else {
    for (int i = lane; i < n; i += 64) {
      for (int d = 0; d < 3; d++) {
        accumulate(frc[d][start + i], prv[d][start + i]);
      }
    }
  }
  ff = wave_sum64(ff);
  gg = wave_sum64(gg);
  dgg = wave_sum64(dgg);
  if (!cg_continue(k, s, lane, n, ff, &step)) {
    return;
  }
  const double gam = (step >= sd_cycles) ? (dgg / gg) : 0.0;
  const bool use_gam = fabs(gam) > 1.0e-12;
  double msum = 0.0;
  // New direction of one component: the previous force is recorded, the conjugate direction
  // replaces the force (when gamma is not negligible) and msum collects its square.
  auto direct = [&](long long ifv, long long &pv, long long &cv, long long &fv) {
    pv = ifv;
    if (use_gam) {
      const double dfcg = static_cast<double>(ifv) + (gam * static_cast<double>(cv));
      msum += dfcg * dfcg;
      const long long v = static_cast<long long>(__builtin_rint(dfcg));
      cv = v;
      fv = v;
    }
    else {
      msum += static_cast<double>(ifv) * static_cast<double>(ifv);
      cv = ifv;
    }
  };
  if (cached) {
#pragma unroll
    for (int it = 0; it < kCgCached; it++) {
      if (lane + (64 * it) < n) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
          direct(rf[it][d], rp[it][d], rc[it][d], rf[it][d]);
        }
      }
    }
  }
  else {
    for (int i = lane; i < n; i += 64) {
      const int g = start + i;
      for (int d = 0; d < 3; d++) {
        long long pv, cv = cg[d][g], fv = frc[d][g];
        direct(fv, pv, cv, fv);
        prv[d][g] = pv;
        cg[d][g] = cv;
        frc[d][g] = fv;
      }
    }
  }
  msum = wave_sum64(msum);
  const double inv_factor = 1.0 / (sqrt(msum) * k.inv_frc_scale);
  auto normalized = [&](long long fv) {
    const double v = static_cast<double>(fv) * inv_factor;
    return k.dp_mode ? static_cast<long long>(__builtin_rint(v)) : static_cast<long long>(v);
  };
  if (cached) {
#pragma unroll
    for (int it = 0; it < kCgCached; it++) {
      const int i = lane + (64 * it);
      if (i < n) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
          prv[d][start + i] = rp[it][d];
          cg[d][start + i] = rc[it][d];
          frc[d][start + i] = normalized(rf[it][d]);
        }
      }
    }
  }
  else {
    for (int i = lane; i < n; i += 64) {
      const int g = start + i;
      for (int d = 0; d < 3; d++) {
        frc[d][g] = normalized(frc[d][g]);
      }
    }
  }
}

hipError_t launch_conjugate_gradient(const CgKit &k, int step, int sd_cycles,
                                     hipStream_t stream) {
  if (k.nsys == 0) {
    return hipSuccess;
  }
  hipLaunchKernelGGL(k_conjugate_gradient, dim3(k.nsys), dim3(64), 0, stream, k, step,
                     sd_cycles);
  return hipGetLastError();
}

// Samples first + steps[s] / every + 1 .. first + count - 1 of system s (the ones it did not
// reach after stopping) repeat its last snapshot.
__global__ void __launch_bounds__(64) k_history_fill(HistKit h, const int *steps, int count) {
  const int s = blockIdx.x;
  const int last = h.first + steps[s] / h.every;
  const int end = min(h.first + count, h.capacity);
  for (int j = last + 1; j < end; j++) {
    const long long *row = h.rows + static_cast<long long>(last) * h.sample_stride +
                           static_cast<long long>(s) * h.sys_stride;
    // Slots 28 and 29 of the row already hold the sums, which write_snapshot recomputes from
    // slots 0..12 and 14 unchanged.
    write_snapshot(h, s, j, row, threadIdx.x);
  }
}

hipError_t launch_history_fill(const HistKit &h, const int *steps, int nsys, int count,
                               hipStream_t stream) {
  if (nsys == 0 || h.rows == nullptr || count <= 0) {
    return hipSuccess;
  }
  hipLaunchKernelGGL(k_history_fill, dim3(nsys), dim3(64), 0, stream, h, steps, count);
  return hipGetLastError();
}

// Line-search state of system s.
__device__ __forceinline__ LineState load_line(const LineKit &k, int s) {
  LineState ls;
  ls.l_move = k.l_move[s];
  ls.s_move = k.s_move[s];
  ls.mfac_a = k.mfac_a[s];
  ls.mfac_b = k.mfac_b[s];
  ls.mfac_c = k.mfac_c[s];
  ls.nrg_a = k.nrg_a[s];
  ls.nrg_b = k.nrg_b[s];
  ls.nrg_c = k.nrg_c[s];
  ls.nrg_d = k.nrg_d[s];
  return ls;
}
__device__ __forceinline__ void store_line(const LineKit &k, int s, const LineState &ls) {
  k.l_move[s] = ls.l_move;
  k.s_move[s] = ls.s_move;
  k.mfac_a[s] = ls.mfac_a;
  k.mfac_b[s] = ls.mfac_b;
  k.mfac_c[s] = ls.mfac_c;
  k.nrg_a[s] = ls.nrg_a;
  k.nrg_b[s] = ls.nrg_b;
  k.nrg_c[s] = ls.nrg_c;
  k.nrg_d[s] = ls.nrg_d;
}

// Line move `move` of system s at potential energy `cur` (src/MolecularMechanics/
// line_movement.cui:67-330): records the energy, updates the state and returns the move in
// fixed-precision coordinate units per fixed-precision direction unit (0 for no move).
__device__ __forceinline__ double line_step(const LineKit &k, int s, int move, double cur,
                                            LineState ls) {
  double mf = line_move_factor(ls, move, cur);
  store_line(k, s, ls);
  if (fabs(mf) > 1.0e-12) {
    mf *= (k.inv_frc_scale * k.gpos_scale) * ((move == 0) ? ls.l_move : ls.s_move);
  }
  else {
    mf = 0.0;
  }
  return mf;
}

// x += mf * direction over a system's atoms, one wavefront (line_movement.cui:343-406).
__device__ __forceinline__ void line_apply(const LineKit &k, int start, int n, double mf,
                                           int lane) {
  if (k.xcrd_ovrf == nullptr && k.xfrc_ovrf == nullptr) {
    for (int i = lane; i < n; i += 64) {
      const int g = start + i;
      k.xcrd[g] += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.xfrc[g])));
      k.ycrd[g] += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.yfrc[g])));
      k.zcrd[g] += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.zfrc[g])));
    }
    return;
  }
  // 95-bit coordinates and / or directions (the overflow forms): the increment is split as
  // doubleToInt95 does and added with the carry of splitFPSum.
  long long *crd[3] = {k.xcrd, k.ycrd, k.zcrd};
  int *crd_o[3] = {k.xcrd_ovrf, k.ycrd_ovrf, k.zcrd_ovrf};
  const long long *frc[3] = {k.xfrc, k.yfrc, k.zfrc};
  const int *frc_o[3] = {k.xfrc_ovrf, k.yfrc_ovrf, k.zfrc_ovrf};
  for (int i = lane; i < n; i += 64) {
    const int g = start + i;
    for (int d = 0; d < 3; d++) {
      const double f = (frc_o[d] != nullptr) ? i95_to_double(frc[d][g], frc_o[d][g])
                                             : static_cast<double>(frc[d][g]);
      long long ix;
      int iy;
      i95_from_double(mf * f, &ix, &iy);
      if (crd_o[d] != nullptr) {
        long long rx;
        int ry;
        i95_add(ix, iy, crd[d][g], crd_o[d][g], &rx, &ry);
        crd[d][g] = rx;
        crd_o[d][g] = ry;
      }
      else {
        crd[d][g] += ix;
      }
    }
  }
}

// One line move per system, src/MolecularMechanics/line_movement.cui:4-415, one wavefront per
// system and four systems per block.  Stopped systems (serving) keep their coordinates and line
// state.
constexpr int kLineWaves = 4;

__global__ void __launch_bounds__(64 * kLineWaves) k_line_advance(LineKit k, int move_number) {
  const int lane = threadIdx.x & 63;
  const int e = (blockIdx.x * kLineWaves) + static_cast<int>(threadIdx.x >> 6);
  if (e >= k.nsys) {
    return;
  }
  const int s = (k.slots != nullptr) ? k.slots[e] : e;
  if (k.done != nullptr && k.done[s] != 0) {
    return;
  }
  // Potential energy: the 13 ScoreCard components, one per lane, summed exactly.
  long long acc = (lane <= 12) ? k.scorecard[(static_cast<size_t>(s) * kScoreStride) + lane] : 0;
  for (int m = 32; m >= 1; m >>= 1) {
    acc += __shfl_xor(acc, m, 64);
  }
  double mf = 0.0;
  if (lane == 0) {
    mf = line_step(k, s, move_number, static_cast<double>(acc) * k.inv_nrg_scale,
                   load_line(k, s));
  }
  mf = __shfl(mf, 0, 64);
  if (mf == 0.0) {
    return;
  }
  line_apply(k, k.sys_start[s], k.sys_natom[s], mf, lane);
}

hipError_t launch_line_advance(const LineKit &k, int move_number, hipStream_t stream) {
  if (k.nsys == 0) {
    return hipSuccess;
  }
  hipLaunchKernelGGL(k_line_advance, dim3((k.nsys + kLineWaves - 1) / kLineWaves),
                     dim3(64 * kLineWaves), 0, stream, k, move_number);
  return hipGetLastError();
}

}  // namespace stormm_amd

namespace stormm_amd {

// ------------------------------------------------------------------------------------------
// Fused minimization cycle.  The reference moves the coordinates after each evaluation with a
// launch of its own (line_movement.cui, launched four times per cycle from
// hpc_minimization.cu:262-405).  Here the move that follows an evaluation is made by the next
// evaluation while it loads the coordinates: its block takes each of its systems' move factor
// from the previous evaluation's ScoreCard row and the line-search state, and adds the move to
// each coordinate as it reads it, writing the moved coordinate back.  A cycle is then the force
// evaluation, the conjugate-gradient kernel and three energy evaluations; the last move of the
// cycle is made by the next cycle's force evaluation (or the final energy evaluation).  The
// arithmetic of each move is k_line_advance's, so trajectories equal the per-step launches' bit
// for bit.
// ------------------------------------------------------------------------------------------

// FORCES selects the force + energy evaluation, otherwise energy only.  pending_move (0..3) is
// the line move to make while loading the coordinates, -1 none; a force evaluation makes move 3
// only for systems that have completed a cycle (so one captured cycle serves the first cycle
// too).  A unit whose systems have all stopped returns at once: its coordinates, forces and
// ScoreCard rows already hold what an evaluation would write.
template <typename T, bool FORCES, bool CLASH>
__global__ void __launch_bounds__(EvalBlockLimit<T>::value)
    k_min_cycle(MMKit k, CycleKit ck, int pending_move) {
  extern __shared__ __align__(16) unsigned char smem[];
  const WuDesc &wu = k.wu[blockIdx.x];
  const int tid = static_cast<int>(threadIdx.x);
  const bool has_sys = tid < wu.sys_count;
  const int s = has_sys ? k.wu_sys[wu.sys_first + tid].sys_index : 0;
  const bool stopped = !has_sys || (ck.cg.done[s] != 0);
  // The pending move's inputs (ScoreCard row and line-search state of the thread's system) are
  // loaded here, so that their latency overlaps the unit's set-up; a stopped system does not
  // move, nor does any system in the force evaluation of its first cycle.
  const bool moves = (pending_move >= 0) && !stopped && !(FORCES && ck.cg.steps[s] == 0);
  LineState ls{};
  long long acc = 0;
  if (moves) {
    ls = load_line(ck.line, s);
    for (int e = 0; e < kNrgSlots; e++) {
      acc += k.scorecard[(static_cast<size_t>(s) * kScoreStride) + e];
    }
  }
  if (__syncthreads_and(stopped ? 1 : 0) != 0) {
    return;
  }
  const int terms = kTermNonbonded | kTermValence;
  StageCtx<T> c(smem, k.maxa, static_cast<int>(blockDim.x >> 6));
  init_ctx<T>(c, k, blockIdx.x, terms);
  stage_load_params<T>(c, k);
  if (pending_move >= 0) {
    // Move factors of the unit's systems (line_step as in k_line_advance, on the exact sum of
    // the 13 ScoreCard components).  The coordinate stage reads them after its first barrier.
    if (has_sys) {
      c.L.move[tid] = moves ? line_step(ck.line, s, pending_move,
                                        static_cast<double>(acc) * ck.line.inv_nrg_scale, ls)
                            : 0.0;
    }
    c.move_pending = true;
  }
  evaluate_wu<T, FORCES, CLASH, false>(c, k, terms);
  stage_store<T, FORCES>(c, k, true);
}

template <typename T, bool FORCES, bool CLASH>
static hipError_t launch_cycle_t(const MMKit &k, const CycleKit &ck, int pending_move,
                                 hipStream_t stream) {
  const int block = std::min(k.block, EvalBlockLimit<T>::value);
  const size_t lds = WuLds<T>::bytes(k.maxa, block / 64);
  auto fn = k_min_cycle<T, FORCES, CLASH>;
  hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(lds));
  if (err != hipSuccess) {
    return err;
  }
  hipLaunchKernelGGL(fn, dim3(k.nwu), dim3(block), lds, stream, k, ck, pending_move);
  return hipGetLastError();
}

hipError_t launch_min_cycle(int precision, const MMKit &k, const CycleKit &ck, bool forces,
                            int pending_move, hipStream_t stream) {
  if (k.nwu == 0) {
    return hipSuccess;
  }
  const bool clash = (k.clash_variant != 0);
  if (precision == 0) {
    if (forces) {
      return clash ? launch_cycle_t<double, true, true>(k, ck, pending_move, stream) :
                     launch_cycle_t<double, true, false>(k, ck, pending_move, stream);
    }
    return clash ? launch_cycle_t<double, false, true>(k, ck, pending_move, stream) :
                   launch_cycle_t<double, false, false>(k, ck, pending_move, stream);
  }
  if (forces) {
    return clash ? launch_cycle_t<float, true, true>(k, ck, pending_move, stream) :
                   launch_cycle_t<float, true, false>(k, ck, pending_move, stream);
  }
  return clash ? launch_cycle_t<float, false, true>(k, ck, pending_move, stream) :
                 launch_cycle_t<float, false, false>(k, ck, pending_move, stream);
}

// ------------------------------------------------------------------------------------------
// State of the allocation-free minimization entry point: per-system cycle counts (int), stop
// flags (int) and RMS forces (double) for up to kNoAllocMaxSystems systems, 16 bytes each.  Each
// device holds its own copy from the library's load until the process ends.  Host code must not
// pass the array's name to a kernel: it stands for the host-side shadow, not the device memory.
// noalloc_scratch() returns the device address.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(void **ptr) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) {
    return err;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    return hipErrorInvalidDevice;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    // Looked up once per device, with that device current.
    void *p = nullptr;
    err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess) {
      return err;
    }
    addr[dev] = p;
  }
  *ptr = addr[dev];
  return hipSuccess;
}

__global__ void __launch_bounds__(256) k_pool_init_slots(SlotInitKit k) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k.count) {
    return;
  }
  const int s = k.slots[e];
  k.slot_start[s] = k.starts[e];
  k.slot_natom[s] = k.natoms[e];
  k.steps[s] = 0;
  k.done[s] = 0;
  k.rms[s] = 0.0;
  for (int v = 0; v < 9; v++) {
    k.line[static_cast<size_t>(v) * k.cap_slots + s] = (v == 0) ? k.dx0 : 0.0;
  }
  for (int c = 0; c < kScoreStride; c++) {
    k.scorecard[static_cast<size_t>(s) * kScoreStride + c] = 0;
  }
}

hipError_t launch_pool_init_slots(const SlotInitKit &k, hipStream_t stream) {
  if (k.count == 0) {
    return hipSuccess;
  }
  hipLaunchKernelGGL(k_pool_init_slots, dim3((k.count + 255) / 256), dim3(256), 0, stream, k);
  return hipGetLastError();
}

}  // namespace stormm_amd
