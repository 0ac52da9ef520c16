// Line-minimization kernels: the conjugate-gradient direction update and the four line moves of
// the reference's launch-per-step minimizer, plus the persistent minimizer that runs every cycle
// of a work unit's systems inside one block.
#include <algorithm>

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

__global__ void __launch_bounds__(64) k_conjugate_gradient(CgKit k, int step, int sd_cycles) {
  const int s = (k.slots != nullptr) ? k.slots[blockIdx.x] : static_cast<int>(blockIdx.x);
  const int lane = threadIdx.x;
  if (k.done != nullptr && k.done[s] != 0) {
    return;
  }
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
    const double cur = static_cast<double>(acc) * k.inv_nrg_scale;
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
    mf = line_move_factor(ls, move_number, cur);
    k.l_move[s] = ls.l_move;
    k.s_move[s] = ls.s_move;
    k.mfac_a[s] = ls.mfac_a;
    k.mfac_b[s] = ls.mfac_b;
    k.mfac_c[s] = ls.mfac_c;
    k.nrg_a[s] = ls.nrg_a;
    k.nrg_b[s] = ls.nrg_b;
    k.nrg_c[s] = ls.nrg_c;
    k.nrg_d[s] = ls.nrg_d;
    if (fabs(mf) > 1.0e-12) {
      mf *= (k.inv_frc_scale * k.gpos_scale) * ((move_number == 0) ? ls.l_move : ls.s_move);
    }
    else {
      mf = 0.0;
    }
  }
  mf = __shfl(mf, 0, 64);
  if (mf == 0.0) {
    return;
  }
  const int start = k.sys_start[s];
  const int n = k.sys_natom[s];
  if (k.xcrd_ovrf == nullptr && k.xfrc_ovrf == nullptr) {
    for (int i = lane; i < n; i += 64) {
      const int g = start + i;
      k.xcrd[g] += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.xfrc[g])));
      k.ycrd[g] += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.yfrc[g])));
      k.zcrd[g] += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.zfrc[g])));
    }
    return;
  }
  // 95-bit coordinates and / or directions (line_movement.cui:343-406, the overflow forms): the
  // increment is split as doubleToInt95 does and added with the carry of splitFPSum.
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
// Persistent minimizer.  One block owns one work unit for the whole minimization: parameters are
// staged in LDS once, and every cycle runs the force + energy evaluation, the conjugate-gradient
// direction update (one wave per system, deterministic butterfly reductions), and the four line
// moves with three energy-only evaluations, exactly the sequence of the reference's host loop
// (src/MolecularMechanics/hpc_minimization.cu:262-405) without returning to the host.
// Coordinates stay in the global fixed-precision arrays (read back into LDS at each evaluation);
// the normalized move direction is kept in the global force arrays as the reference does.  The
// host launches it only for syntheses without restraints, so its evaluations leave the restraint
// stage out.
// ------------------------------------------------------------------------------------------
struct MinLds {
  LineState ls[kMaxWuSystems];
  double mf[kMaxWuSystems];
  double rms[kMaxWuSystems];
  int done[kMaxWuSystems];
  int cycles[kMaxWuSystems];
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int m = 32; m >= 1; m >>= 1) {
    v += __shfl_xor(v, m, 64);
  }
  return v;
}

template <typename T>
__device__ __forceinline__ void apply_moves(StageCtx<T> &c, const MMKit &k, const MinLds *M) {
  for (int a = c.tid; a < c.wu.lds_atoms; a += c.nthr) {
    const int g = c.L.gidx[a];
    if (g < 0) {
      continue;
    }
    const double mf = M->mf[c.L.slot[a]];
    if (mf != 0.0) {
      k.xcrd[g] += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.xfrc[g])));
      k.ycrd[g] += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.yfrc[g])));
      k.zcrd[g] += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.zfrc[g])));
    }
  }
  __syncthreads();
}

template <typename T>
__device__ double wu_energy(const StageCtx<T> &c, int s) {
  long long acc = 0;
  for (int e = 0; e <= 12; e++) {
    acc += c.L.nrg[s * 16 + e];
  }
  return static_cast<double>(acc);
}

// CLASH instantiations serve minimizations with clash-damping cycles: every evaluation takes
// each system's damping from the cycle counter (off once the damped cycles are over, which gives
// the plain potentials), so one inlined evaluation of each kind serves the whole run.
template <typename T, bool CLASH>
__global__ void __launch_bounds__(kMaxMinimizeBlock) k_minimize(MinKit mk) {
  extern __shared__ __align__(16) unsigned char smem[];
  const MMKit &k = mk.mm;
  const int waves = static_cast<int>(blockDim.x >> 6);
  StageCtx<T> c(smem, k.maxa, waves);
  MinLds *M = reinterpret_cast<MinLds *>(smem + WuLds<T>::bytes(k.maxa, waves));
  const int terms = kTermNonbonded | kTermValence;
  init_ctx<T>(c, k, blockIdx.x, terms);
  stage_load_params<T>(c, k);
  const int nsys = c.wu.sys_count;
  const double inv_frc = ldexp(1.0, -k.frc_bits);
  const double inv_nrg = ldexp(1.0, -k.nrg_bits);
  const double move_scale = inv_frc * ldexp(1.0, k.gpos_bits);
  const bool dp_mode = (sizeof(T) == 8);
  if (c.tid < nsys) {
    LineState ls;
    ls.l_move = mk.dx0;
    ls.s_move = mk.dx0;
    ls.mfac_a = ls.mfac_b = ls.mfac_c = 0.0;
    ls.nrg_a = ls.nrg_b = ls.nrg_c = ls.nrg_d = 0.0;
    M->ls[c.tid] = ls;
    M->mf[c.tid] = 0.0;
    M->rms[c.tid] = 0.0;
    M->done[c.tid] = 0;
    M->cycles[c.tid] = 0;
  }
  // PhaseSpaceSynthesis::primeConjugateGradientCalculation for this unit's atoms.
  for (int a = c.tid; a < c.wu.lds_atoms; a += c.nthr) {
    const int g = c.L.gidx[a];
    if (g >= 0) {
      mk.xprv[g] = 0;
      mk.yprv[g] = 0;
      mk.zprv[g] = 0;
      mk.xcg[g] = 0;
      mk.ycg[g] = 0;
      mk.zcg[g] = 0;
    }
  }
  __syncthreads();
  const int wave = c.tid >> 6;
  const int lane = c.tid & 63;
  const int nwaves = c.nthr >> 6;
  for (int cyc = 0; cyc < mk.total_cycles; cyc++) {
    bool all_done = true;
    for (int s = 0; s < nsys; s++) {
      all_done = all_done && (M->done[s] != 0);
    }
    if (all_done) {
      break;
    }
    c.cycle = cyc;
    evaluate_wu<T, true, CLASH, false>(c, k, terms);

    // Convergence test and conjugate-gradient direction, one wave per system.  The three
    // Cartesian components are handled by the same code on their own arrays.
    for (int s = wave; s < nsys; s += nwaves) {
      const WuSys ws = c.L.sys[s];
      double ff = 0.0, gg = 0.0, dgg = 0.0;
      auto sums = [&](const long long *lf, const long long *pv) {
        for (int i = lane; i < ws.natom; i += 64) {
          const long long lfv = lf[ws.lds_base + i];
          const long long lpv = pv[ws.global_start + i];
          const double df = static_cast<double>(lfv);
          const double dp = static_cast<double>(lpv);
          ff += df * df;
          gg += dp * dp;
          dgg += static_cast<double>(lfv - lpv) * df;
        }
      };
      sums(c.L.fx, mk.xprv);
      sums(c.L.fy, mk.yprv);
      sums(c.L.fz, mk.zprv);
      ff = wave_sum(ff);
      gg = wave_sum(gg);
      dgg = wave_sum(dgg);
      const double rms = sqrt(ff / (3.0 * ws.natom)) * inv_frc;
      const bool done = (M->done[s] != 0) || (mk.conv_rms > 0.0 && rms <= mk.conv_rms);
      double mf0 = 0.0;
      if (!done) {
        const int step = cyc + 1;
        const double gam = (step >= mk.sd_cycles) ? (dgg / gg) : 0.0;
        const bool use_gam = fabs(gam) > 1.0e-12;
        double msum = 0.0;
        auto direction = [&](const long long *lf, long long *pv, long long *cg, long long *gf) {
          for (int i = lane; i < ws.natom; i += 64) {
            const int g = ws.global_start + i;
            const long long ifv = lf[ws.lds_base + i];
            pv[g] = ifv;
            if (use_gam) {
              const double dfcg = static_cast<double>(ifv) + (gam * static_cast<double>(cg[g]));
              msum += dfcg * dfcg;
              const long long v = static_cast<long long>(__builtin_rint(dfcg));
              cg[g] = v;
              gf[g] = v;
            }
            else {
              msum += static_cast<double>(ifv) * static_cast<double>(ifv);
              cg[g] = ifv;
              gf[g] = ifv;
            }
          }
        };
        direction(c.L.fx, mk.xprv, mk.xcg, k.xfrc);
        direction(c.L.fy, mk.yprv, mk.ycg, k.yfrc);
        direction(c.L.fz, mk.zprv, mk.zcg, k.zfrc);
        msum = wave_sum(msum);
        const double inv_factor = 1.0 / (sqrt(msum) * inv_frc);
        auto normalize = [&](long long *gf) {
          for (int i = lane; i < ws.natom; i += 64) {
            const int g = ws.global_start + i;
            const double v = static_cast<double>(gf[g]) * inv_factor;
            gf[g] = dp_mode ? static_cast<long long>(__builtin_rint(v)) :
                              static_cast<long long>(v);
          }
        };
        normalize(k.xfrc);
        normalize(k.yfrc);
        normalize(k.zfrc);
      }
      if (lane == 0) {
        M->rms[s] = rms;
        if (!done) {
          LineState ls = M->ls[s];
          mf0 = line_move_factor(ls, 0, wu_energy(c, s) * inv_nrg);
          M->ls[s] = ls;
          M->mf[s] = (fabs(mf0) > 1.0e-12) ? mf0 * (move_scale * ls.l_move) : 0.0;
          M->cycles[s] = cyc + 1;
        }
        else {
          M->done[s] = 1;
          M->mf[s] = 0.0;
        }
      }
    }
    __syncthreads();
    apply_moves<T>(c, k, M);
    for (int move = 1; move < 4; move++) {
      evaluate_wu<T, false, CLASH, false>(c, k, terms);
      for (int s = c.tid; s < nsys; s += c.nthr) {
        if (M->done[s] == 0) {
          LineState ls = M->ls[s];
          const double mf = line_move_factor(ls, move, wu_energy(c, s) * inv_nrg);
          M->ls[s] = ls;
          M->mf[s] = (fabs(mf) > 1.0e-12) ? mf * (move_scale * ls.s_move) : 0.0;
        }
        else {
          M->mf[s] = 0.0;
        }
      }
      __syncthreads();
      apply_moves<T>(c, k, M);
    }
  }
  // Final energies of every system (hpc_minimization.cu:395-405), undamped.
  c.cycle = k.clash_cycles;
  evaluate_wu<T, false, CLASH, false>(c, k, terms);
  stage_store<T, false>(c, k, true);
  for (int s = c.tid; s < nsys; s += c.nthr) {
    const int si = c.L.sys[s].sys_index;
    if (mk.cycles_run != nullptr) {
      mk.cycles_run[si] = M->cycles[s];
    }
    if (mk.rms_force != nullptr) {
      mk.rms_force[si] = M->rms[s];
    }
    if (mk.l_move != nullptr) {
      mk.l_move[si] = M->ls[s].l_move;
    }
  }
}

template <typename T>
static hipError_t launch_min_t(const MinKit &k, hipStream_t stream) {
  const int block = std::min(k.mm.block, kMaxMinimizeBlock);
  const size_t lds = WuLds<T>::bytes(k.mm.maxa, block / 64) + sizeof(MinLds);
  auto fn = (k.mm.clash_variant != 0) ? k_minimize<T, true> : k_minimize<T, false>;
  hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(lds));
  if (err != hipSuccess) {
    return err;
  }
  hipLaunchKernelGGL(fn, dim3(k.mm.nwu), dim3(block), lds, stream, k);
  return hipGetLastError();
}

hipError_t launch_minimize_persistent(int precision, const MinKit &k, hipStream_t stream) {
  if (k.mm.nwu == 0) {
    return hipSuccess;
  }
  return (precision == 0) ? launch_min_t<double>(k, stream) : launch_min_t<float>(k, stream);
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
