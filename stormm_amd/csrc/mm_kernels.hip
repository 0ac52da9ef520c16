// Fused batched energy / force evaluation kernels and their launchers.
#include <algorithm>

#include "mm_eval.h"
#include "kernels.h"

namespace stormm_amd {

// CLASH instantiations carry the clash-damped Coulomb / Lennard-Jones / 1:4 potentials used in
// the first cycles of a minimization; the others never pay for them.
template <typename T, bool FORCES, bool CLASH>
__global__ void __launch_bounds__(EvalBlockLimit<T>::value)
    k_mm_eval(MMKit k, int terms, int eval_energy) {
  extern __shared__ __align__(16) unsigned char smem[];
  StageCtx<T> c(smem, k.maxa, static_cast<int>(blockDim.x >> 6));
  init_ctx<T>(c, k, blockIdx.x, terms);
  stage_load_params<T>(c, k);
  evaluate_wu<T, FORCES, CLASH>(c, k, terms);
  stage_store<T, FORCES>(c, k, eval_energy != 0);
}

template <typename T, bool FORCES, bool CLASH>
static hipError_t launch_eval_t(const MMKit &k, int terms, int eval_energy, hipStream_t stream) {
  const int block = std::min(k.block, EvalBlockLimit<T>::value);
  const size_t lds = WuLds<T>::bytes(k.maxa, block / 64);
  auto fn = k_mm_eval<T, FORCES, CLASH>;
  hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(lds));
  if (err != hipSuccess) {
    return err;
  }
  hipLaunchKernelGGL(fn, dim3(k.nwu), dim3(block), lds, stream, k, terms, eval_energy);
  return hipGetLastError();
}

hipError_t launch_mm_eval(int precision, const MMKit &k, int terms, int eval_force,
                          int eval_energy, hipStream_t stream) {
  if (k.nwu == 0) {
    return hipSuccess;
  }
  const bool clash = (k.clash_variant != 0);
  if (precision == 0) {
    if (clash) {
      return eval_force ? launch_eval_t<double, true, true>(k, terms, eval_energy, stream) :
                          launch_eval_t<double, false, true>(k, terms, eval_energy, stream);
    }
    return eval_force ? launch_eval_t<double, true, false>(k, terms, eval_energy, stream) :
                        launch_eval_t<double, false, false>(k, terms, eval_energy, stream);
  }
  if (clash) {
    return eval_force ? launch_eval_t<float, true, true>(k, terms, eval_energy, stream) :
                        launch_eval_t<float, false, true>(k, terms, eval_energy, stream);
  }
  return eval_force ? launch_eval_t<float, true, false>(k, terms, eval_energy, stream) :
                      launch_eval_t<float, false, false>(k, terms, eval_energy, stream);
}

// ------------------------------------------------------------------------------------------
// Zeroing kernels: PhaseSpaceSynthesis::initializeForces (kPsyInitializeForces,
// src/Synthesis/hpc_phasespace_synthesis.cu:243) and ScoreCard::initialize
// (kInitializeScoreCard, src/Potential/hpc_scorecard.cu:17).  Plain fill kernels keep these
// graph-capturable.
__global__ void k_fill_ll(long long *a, long long *b, long long *c, size_t n, long long v) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    a[i] = v;
    if (b != nullptr) {
      b[i] = v;
    }
    if (c != nullptr) {
      c[i] = v;
    }
  }
}

hipError_t launch_fill(long long *a, long long *b, long long *c, size_t n, hipStream_t stream) {
  if (n == 0) {
    return hipSuccess;
  }
  const int block = 256;
  const size_t want = (n + block - 1) / block;
  const int grid = static_cast<int>(want < 4096 ? want : 4096);
  hipLaunchKernelGGL(k_fill_ll, dim3(grid), dim3(block), 0, stream, a, b, c, n, 0LL);
  return hipGetLastError();
}

__global__ void k_fill_int(int *a, int *b, int *c, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    a[i] = 0;
    b[i] = 0;
    c[i] = 0;
  }
}

hipError_t launch_fill_int(int *a, int *b, int *c, size_t n, hipStream_t stream) {
  if (n == 0) {
    return hipSuccess;
  }
  const int block = 256;
  const size_t want = (n + block - 1) / block;
  const int grid = static_cast<int>(want < 4096 ? want : 4096);
  hipLaunchKernelGGL(k_fill_int, dim3(grid), dim3(block), 0, stream, a, b, c, n);
  return hipGetLastError();
}

// ScoreCard::commit (kScoreCardCommit, src/Potential/hpc_scorecard.cu:170-215) with the
// potential and total energy sums of kScoreCardSumEnergy (:72-157): fold the instantaneous
// accumulators into running sums and sums of squares and, when series is given, store them as
// this sample's time-series row, the potential energy (BOND..RESTRAINT) in slot 28 and the total
// (potential + KINETIC, slot 14) in slot 29.
__global__ void k_scorecard_commit(const long long *inst, double *running, double *squared,
                                   long long *series, int nsys, double inv_scale) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) {
    return;
  }
  const long long *row = inst + static_cast<size_t>(s) * kScoreStride;
  long long pe = 0;
  for (int k = 0; k < kNrgSlots; k++) {
    pe += row[k];
  }
  for (int k = 0; k < kScoreStride; k++) {
    const long long v = (k == kPotentialSlot) ? pe : (k == kTotalSlot) ? pe + row[kKineticSlot] :
                                                                         row[k];
    const double d = static_cast<double>(v) * inv_scale;
    running[static_cast<size_t>(s) * kScoreStride + k] += d;
    squared[static_cast<size_t>(s) * kScoreStride + k] += d * d;
    if (series != nullptr) {
      series[static_cast<size_t>(s) * kScoreStride + k] = v;
    }
  }
}

hipError_t launch_scorecard_commit(const long long *inst, double *running, double *squared,
                                   long long *series, int nsys, double inv_scale,
                                   hipStream_t stream) {
  if (nsys == 0) {
    return hipSuccess;
  }
  hipLaunchKernelGGL(k_scorecard_commit, dim3((nsys + 127) / 128), dim3(128), 0, stream, inst,
                     running, squared, series, nsys, inv_scale);
  return hipGetLastError();
}

}  // namespace stormm_amd
