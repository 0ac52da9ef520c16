// Device-side arithmetic shared by the fused evaluation kernel and the persistent minimizer:
// precision traits, DPP row rotation, the Born-radius integral and its radial derivative, the
// work-unit LDS layout and the valence term evaluators.
#pragma once

#include "int95.h"
#include "internal.h"

namespace stormm_amd {

// Generalized Born Taylor-series coefficients (far-field expansion of the HCT integral),
// src/Constants/generalized_born.h:27-36.
constexpr double kGTA = 0.33333333333333333333, kGTB = 0.4, kGTC = 0.42857142857142857143,
                 kGTD = 0.44444444444444444444, kGTDD = 0.45454545454545454545;
constexpr double kGTE = 1.33333333333333333333, kGTF = 2.4, kGTG = 3.42857142857142857143,
                 kGTH = 4.44444444444444444444, kGTHH = 5.45454545454545454545;
constexpr double kGbKscale = 0.73;      // src/Constants/generalized_born.h:120
constexpr double kGbOffset = 0.09;      // src/Constants/generalized_born.h:112

// ------------------------------------------------------------------------------------------
// Precision traits.  Single precision uses the hardware transcendental units (v_rsq, v_rcp,
// v_log, v_exp); double precision uses the correctly rounded library routines.
// ------------------------------------------------------------------------------------------
// Nearest integer (ties to even, as rint) of a double as int64.  For |x| < 2^51, adding
// 1.5 * 2^52 rounds x to an integer held in the low bits of the sum's bit pattern; subtracting
// the constant's pattern recovers it in two instructions.  Larger or non-finite values take the
// library conversion.
__device__ __forceinline__ long long round_to_ll(double x) {
  // x must be rounded to double before the addition: never fuse it with the product that made x.
#pragma clang fp contract(off)
  constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
  if (__builtin_expect(fabs(x) < 2251799813685248.0, 1)) {
    return __double_as_longlong(x + kMagic) - __double_as_longlong(kMagic);
  }
  return static_cast<long long>(__builtin_rint(x));
}

template <typename T> struct Prec;

template <> struct Prec<float> {
  using crd_t = int;
  static __device__ __forceinline__ float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_sqrtf(x); }
  static __device__ __forceinline__ float cbrt(float x) { return ::cbrtf(x); }
  static __device__ __forceinline__ float log(float x) {
    return __builtin_amdgcn_logf(x) * 0.693147180559945309417f;
  }
  static __device__ __forceinline__ float exp(float x) {
    return __builtin_amdgcn_exp2f(x * 1.44269504088896340736f);
  }
  static __device__ __forceinline__ long long fixed(float x) {
    return round_to_ll(static_cast<double>(x));
  }
};

template <> struct Prec<double> {
  using crd_t = double;
  // The hardware estimates refined by two Newton-Raphson steps, y <- y + y (1 - x y) for the
  // reciprocal and y <- y + (y / 2) (1 - x y^2) for the reciprocal square root: each step about
  // squares the estimate's relative error, so two reach the last few bits of a double.  For
  // positive, finite, normal arguments (distances and radii here).
  static __device__ __forceinline__ double rsqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
    y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    return fma(0.5 * y, fma(-x * y, y, 1.0), y);
  }
  static __device__ __forceinline__ double rcp(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    return fma(y, fma(-x, y, 1.0), y);
  }
  static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
  static __device__ __forceinline__ double cbrt(double x) { return ::cbrt(x); }
  static __device__ __forceinline__ double log(double x) { return ::log(x); }
  static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
  static __device__ __forceinline__ long long fixed(double x) { return round_to_ll(x); }
};

// ------------------------------------------------------------------------------------------
// Rotation of a value one lane to the right inside each 16-lane DPP row (row_ror:1): after k
// rotations lane L holds what lane (L - k) mod 16 of its row held.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int row_rotate(int v) {
  // Every lane of a row has a source under row_ror, so the bound-control and old-value
  // operands never take effect; mov_dpp leaves the old value undefined, which lets the compiler
  // rotate a register in place instead of zero-filling a fresh one each step.
  return __builtin_amdgcn_mov_dpp(v, 0x121, 0xF, 0xF, true);
}
__device__ __forceinline__ float row_rotate(float v) {
  return __int_as_float(row_rotate(__float_as_int(v)));
}
__device__ __forceinline__ double row_rotate(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = row_rotate(static_cast<int>(bits & 0xffffffffLL));
  const int hi = row_rotate(static_cast<int>(bits >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              (static_cast<long long>(lo) & 0xffffffffLL));
}

// The value that lane (L - K) mod 16 of the row holds, K = 0..15 (row_ror:K): the j-atom data a
// lane meets at rotation step K while the data stay where they were loaded.  With one use per
// rotated value the compiler folds the move into the arithmetic instruction as its DPP source
// operand, so reading rotated data costs no instruction of its own.
template <int K> __device__ __forceinline__ int row_ror(int v) {
  if constexpr (K == 0) {
    return v;
  }
  else {
    return __builtin_amdgcn_mov_dpp(v, 0x120 + K, 0xF, 0xF, true);
  }
}
template <int K> __device__ __forceinline__ float row_ror(float v) {
  return __int_as_float(row_ror<K>(__float_as_int(v)));
}
template <int K> __device__ __forceinline__ double row_ror(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = row_ror<K>(static_cast<int>(bits & 0xffffffffLL));
  const int hi = row_ror<K>(static_cast<int>(bits >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              (static_cast<long long>(lo) & 0xffffffffLL));
}

// Calls f(std::integral_constant<int, K>) for each K of the sequence, in order (rotation steps
// unrolled with compile-time rotation counts).
// A scheduling barrier after each step keeps the steps' live values from overlapping (the
// unrolled steps would otherwise be interleaved past the register budget).
template <typename F, int... K>
__device__ __forceinline__ void unrolled_steps(F &&f, std::integer_sequence<int, K...>) {
  ((f(std::integral_constant<int, K>{}), __builtin_amdgcn_sched_barrier(0)), ...);
}

// Makes x final at this point: an empty statement that reads and rewrites its register, so that
// the arithmetic producing x is not moved below it (into a later block, where the rotated
// operands that arithmetic reads would each need a separate move).
template <typename T> __device__ __forceinline__ void settle(T &x) {
  asm volatile("" : "+v"(x));
}

// Sum of a 64-bit integer over the 16 lanes of a row.
__device__ __forceinline__ long long row_sum(long long v) {
  v += __shfl_xor(v, 8, 16);
  v += __shfl_xor(v, 4, 16);
  v += __shfl_xor(v, 2, 16);
  v += __shfl_xor(v, 1, 16);
  return v;
}

__device__ __forceinline__ void lds_add(long long *addr, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long *>(addr), static_cast<unsigned long long>(v));
}

// ------------------------------------------------------------------------------------------
// Born-radius integral contribution of a sphere of screened radius sj at distance r onto an
// atom of offset radius ri (HCT pairwise descreening), src/Potential/nonbonded_potential.tpp:
// 578-617, and its radial derivative divided by r (sign convention of :1138-1188).  The
// reference takes a far-field Taylor series in (sj / r)^2 when r > 4 sj and one of three
// analytic branches otherwise:
//   b2  r > ri + sj           : -sj / (2 (r^2 - sj^2))             - ln(|r - sj| / (r + sj)) / (4 r)
//   b3  |ri - sj| < r <= ri+sj: -(1/ri (2 - theta) - 1/(r + sj)) / 4 - ln(ri / (r + sj)) / (4 r)
//   b4  r <= |ri - sj|, ri<sj : -sj / (2 (r^2 - sj^2)) - 1/ri       - ln(|r - sj| / (r + sj)) / (4 r)
// (zero when none applies).  The far and near forms are separate functions so that the tile
// loops can evaluate the cheap series for every lane and hand the rarer near pairs to a
// compacted pass.  The far-field series, -s^3 / r^4 (a + b x + c x^2 + d x^3 + dd x^4) with
// x = (s / r)^2, and its derivative follow :595-598 and :1121-1124.  The near branches share
// one reciprocal and one logarithm: b3 needs 1/(r + sj), b2 and b4 need 1/(r + sj) and
// 1/(r^2 - sj^2), both taken from 1/((r + sj)(r - sj)) since |r - sj| >= ri keeps that product
// away from zero in those branches.
// ------------------------------------------------------------------------------------------
// Squared distance given to an absent pair in the Born-integral and derivative tile loops: far
// beyond 4 screened radii of any atom, with 1 / r^4 below the smallest denormal, so both far-field
// terms round to exactly zero and the pair never enters the near-field queue.
// kCrd is the coordinate (on all three axes, +kCrd for an absent i atom, -kCrd for an absent j
// atom) that puts an absent atom of a tile as far away, with or without a partner present.
template <typename T> struct FarAway;
template <> struct FarAway<float> {
  static constexpr float kR2 = 1.0e30f;
  static constexpr float kCrd = 1.0e15f;
};
template <> struct FarAway<double> {
  static constexpr double kR2 = 1.0e200;
  static constexpr double kCrd = 1.0e100;
};

// Far-field series in x = (s / r)^2: the Born-integral term is -s^3 / r^4 times psi_series(x),
// the derivative term (divided by r) s^3 / r^6 times dpsi_series(x).
template <typename T>
__device__ __forceinline__ T psi_series(T x) {
  return T(kGTA) + x * (T(kGTB) + x * (T(kGTC) + x * (T(kGTD) + x * T(kGTDD))));
}

template <typename T>
__device__ __forceinline__ T dpsi_series(T x) {
  return T(kGTE) + x * (T(kGTF) + x * (T(kGTG) + x * (T(kGTH) + x * T(kGTHH))));
}

// Shared set-up of the near branches: branch flags, 1/(r + sj), 1/(r^2 - sj^2) (b2, b4) and
// the logarithm's argument.
template <typename T> struct NearBranch {
  bool b3, any, b4;
  T uij, t1, lg;
  __device__ __forceinline__ NearBranch(T r, T ri, T sj) {
    using P = Prec<T>;
    const bool b2 = r > ri + sj;
    b3 = (!b2) && (r > fabs(ri - sj));
    b4 = (!b2) && (!b3) && (ri < sj);
    any = b2 || b3 || b4;
    const T rps = r + sj;
    const T rms = r - sj;
    const T inv = P::rcp(b3 ? rps : rps * rms);
    uij = b3 ? inv : rms * inv;
    t1 = inv;
    lg = P::log((b3 ? ri : fabs(rms)) * uij);
  }
};

template <typename T>
__device__ __forceinline__ T psi_near(T r, T r2, T invr, T ri, T iri, T sj) {
  const NearBranch<T> nb(r, ri, sj);
  const T theta = T(0.5) * iri * invr * (r2 + (ri * ri) - (sj * sj));
  const T a3 = -T(0.25) * ((iri * (T(2) - theta)) - nb.uij);
  const T a24 = -(T(0.5) * sj * nb.t1) - (nb.b4 ? iri : T(0));
  const T res = (nb.b3 ? a3 : a24) - (T(0.25) * invr * nb.lg);
  return nb.any ? res : T(0);
}

template <typename T>
__device__ __forceinline__ T dpsi_near(T r, T r2, T invr, T invr2, T ri, T iri, T sj) {
  const NearBranch<T> nb(r, ri, sj);
  const T invr3 = invr2 * invr;
  const T a3 = -T(0.25) * ((-T(0.5) * (r2 - (ri * ri) + (sj * sj)) * invr3 * iri * iri) +
                           (invr * nb.uij * (nb.uij - invr)));
  const T a24 = sj * nb.t1 * (nb.t1 - (T(0.5) * invr2));
  const T res = (nb.b3 ? a3 : a24) + (T(0.25) * invr3 * nb.lg);
  return nb.any ? res : T(0);
}

// Fixed-precision image round(x * 2^bits) of a pair contribution added to an LDS accumulator.
// Single precision converts through 32 bits (rounding at 2^-min(bits, 28), then shifting) when
// the value fits, and through double precision otherwise.
__device__ __forceinline__ long long fixed_contrib(float x, int bits) {
  const int lo = (bits < 28) ? bits : 28;
  const float v = x * __builtin_ldexpf(1.0f, lo);
  if (fabsf(v) < 2147483520.0f) {
    return static_cast<long long>(__float2int_rn(v)) << (bits - lo);
  }
  return round_to_ll(ldexp(static_cast<double>(x), bits));
}
__device__ __forceinline__ long long fixed_contrib(double x, int bits) {
  return round_to_ll(ldexp(x, bits));
}

// Largest block a kernel of each precision is compiled for.
template <typename T> struct EvalBlockLimit;
template <> struct EvalBlockLimit<float> {
  static constexpr int value = kMaxEvalBlock;
};
template <> struct EvalBlockLimit<double> {
  static constexpr int value = kMaxEvalBlockDouble;
};

// ------------------------------------------------------------------------------------------
// Work-unit LDS layout.
// ------------------------------------------------------------------------------------------
template <typename T> struct WuLds {
  using crd_t = typename Prec<T>::crd_t;
  int4 *tasks;              // the unit's tile tasks, placed at their systems' LDS bases
  long long *psi, *dedr, *fx, *fy, *fz, *nrg, *org;
  double *move;             // per system: pending line move (coordinate units per direction unit)
  T *q, *p1, *p2, *rho, *irho, *scr, *rad, *irad, *sd;
  crd_t *cx, *cy, *cz;
  int *gidx, *slot;
  WuSys *sys;
  int *cpre, *opre, *dpre;  // per slot: first clean off-diagonal task, first other off-diagonal
                            // task (counted among those), first diagonal task
  int *vpre;                // per kind (bonds, angles, dihedrals): each slot's first term
  int *work;                // per-stage work-item counters (kStageCounters)
  T *clash_d, *clash_r;     // per-system clash distance and sigma ratio of this evaluation
  int *clash_on;            // per-system flag: damped potentials apply
  unsigned *nearq;          // kNearQueueStride entries per wavefront: its near-pair queue

  static __host__ __device__ size_t bytes(int maxa, int waves) {
    const size_t m = static_cast<size_t>(maxa);
    return (kMaxWuTasks * sizeof(int4)) +
           (5 * m * sizeof(long long)) + (kMaxWuSystems * 16 * sizeof(long long)) +
           (3 * kMaxWuSystems * sizeof(long long)) + (kMaxWuSystems * sizeof(double)) +
           (9 * m * sizeof(T)) +
           (3 * m * sizeof(crd_t)) + (2 * m * sizeof(int)) + (kMaxWuSystems * sizeof(WuSys)) +
           (6 * (kMaxWuSystems + 1) * sizeof(int)) + (kStageCounters * sizeof(int)) +
           (2 * kMaxWuSystems * sizeof(T)) + (kMaxWuSystems * sizeof(int)) +
           (static_cast<size_t>(waves) * kNearQueueStride * sizeof(unsigned));
  }

  __device__ WuLds(unsigned char *base, int maxa, int waves) {
    unsigned char *p = base;
    auto take = [&](size_t nbytes) {
      unsigned char *r = p;
      p += nbytes;
      return r;
    };
    const size_t m = static_cast<size_t>(maxa);
    tasks = reinterpret_cast<int4 *>(take(kMaxWuTasks * sizeof(int4)));
    psi = reinterpret_cast<long long *>(take(m * sizeof(long long)));
    dedr = reinterpret_cast<long long *>(take(m * sizeof(long long)));
    fx = reinterpret_cast<long long *>(take(m * sizeof(long long)));
    fy = reinterpret_cast<long long *>(take(m * sizeof(long long)));
    fz = reinterpret_cast<long long *>(take(m * sizeof(long long)));
    nrg = reinterpret_cast<long long *>(take(kMaxWuSystems * 16 * sizeof(long long)));
    org = reinterpret_cast<long long *>(take(3 * kMaxWuSystems * sizeof(long long)));
    move = reinterpret_cast<double *>(take(kMaxWuSystems * sizeof(double)));
    q = reinterpret_cast<T *>(take(m * sizeof(T)));
    p1 = reinterpret_cast<T *>(take(m * sizeof(T)));
    p2 = reinterpret_cast<T *>(take(m * sizeof(T)));
    rho = reinterpret_cast<T *>(take(m * sizeof(T)));
    irho = reinterpret_cast<T *>(take(m * sizeof(T)));
    scr = reinterpret_cast<T *>(take(m * sizeof(T)));
    rad = reinterpret_cast<T *>(take(m * sizeof(T)));
    irad = reinterpret_cast<T *>(take(m * sizeof(T)));
    sd = reinterpret_cast<T *>(take(m * sizeof(T)));
    clash_d = reinterpret_cast<T *>(take(kMaxWuSystems * sizeof(T)));
    clash_r = reinterpret_cast<T *>(take(kMaxWuSystems * sizeof(T)));
    cx = reinterpret_cast<crd_t *>(take(m * sizeof(crd_t)));
    cy = reinterpret_cast<crd_t *>(take(m * sizeof(crd_t)));
    cz = reinterpret_cast<crd_t *>(take(m * sizeof(crd_t)));
    gidx = reinterpret_cast<int *>(take(m * sizeof(int)));
    slot = reinterpret_cast<int *>(take(m * sizeof(int)));
    sys = reinterpret_cast<WuSys *>(take(kMaxWuSystems * sizeof(WuSys)));
    cpre = reinterpret_cast<int *>(take((kMaxWuSystems + 1) * sizeof(int)));
    opre = reinterpret_cast<int *>(take((kMaxWuSystems + 1) * sizeof(int)));
    dpre = reinterpret_cast<int *>(take((kMaxWuSystems + 1) * sizeof(int)));
    vpre = reinterpret_cast<int *>(take(3 * (kMaxWuSystems + 1) * sizeof(int)));
    work = reinterpret_cast<int *>(take(kStageCounters * sizeof(int)));
    clash_on = reinterpret_cast<int *>(take(kMaxWuSystems * sizeof(int)));
    nearq = reinterpret_cast<unsigned *>(take(static_cast<size_t>(waves) * kNearQueueStride *
                                              sizeof(unsigned)));
  }
};

// Displacement between two LDS coordinates in Angstroms.
__device__ __forceinline__ float crd_delta(int a, int b, float scale) {
  return static_cast<float>(a - b) * scale;
}
__device__ __forceinline__ double crd_delta(double a, double b, double) { return a - b; }

// exp(-r^2 / (4 Ri Rj)) of the GB pair function as eval(r^2 * (1 / Rj) * (kScale / Ri)): the
// hardware base-2 exponential in single precision (kScale carries log2 e), the library
// exponential in double precision.
template <typename T> struct GbExp;
template <> struct GbExp<float> {
  static constexpr float kScale = -0.25f * 1.44269504088896340736f;
  static __device__ __forceinline__ float eval(float x) { return __builtin_amdgcn_exp2f(x); }
};
template <> struct GbExp<double> {
  static constexpr double kScale = -0.25;
  static __device__ __forceinline__ double eval(double x) { return ::exp(x); }
};

// Coordinates of a pair tile's atoms in Angstroms in the working precision.  Single precision
// converts the LDS fixed-precision coordinates relative to a point between the tile's two atom
// blocks once per task, as the reference imports tile coordinates relative to the tile's centre
// (src/Potential/nonbonded_potential_tilegroups.cui:395-426), so the rotation steps take plain
// differences.  Double precision holds Angstroms already.
__device__ __forceinline__ int tile_centre(int a, int b) { return (a >> 1) + (b >> 1); }
__device__ __forceinline__ double tile_centre(double, double) { return 0.0; }
__device__ __forceinline__ float tile_crd(int v, int centre, float scale) {
  return static_cast<float>(v - centre) * scale;
}
__device__ __forceinline__ double tile_crd(double v, double, double) { return v; }

// Conversion of a fixed-precision global coordinate to the LDS representation.
template <typename T> struct CrdConv;
template <> struct CrdConv<float> {
  static __device__ __forceinline__ int to_lds(long long v, long long origin, int shift) {
    const long long d = v - origin;
    const long long half = (shift > 0) ? (1LL << (shift - 1)) : 0LL;
    return static_cast<int>((d + half) >> shift);
  }
  static __device__ __forceinline__ bool in_range(long long v, long long origin, int shift) {
    const long long d = (v - origin) >> shift;
    return d > -2147483000LL && d < 2147483000LL;
  }
  // Displacements from a 95-bit value (exact 128-bit difference from the origin).
  static __device__ __forceinline__ int to_lds_wide(__int128 d, int shift) {
    const __int128 half = (shift > 0) ? (static_cast<__int128>(1) << (shift - 1)) : 0;
    return static_cast<int>((d + half) >> shift);
  }
  static __device__ __forceinline__ bool in_range_wide(__int128 d, int shift) {
    const __int128 q = d >> shift;
    return q > -2147483000 && q < 2147483000;
  }
};
template <> struct CrdConv<double> {
  static __device__ __forceinline__ double to_lds(long long v, long long origin, int shift) {
    return static_cast<double>(v - origin) * ldexp(1.0, -shift);
  }
  static __device__ __forceinline__ bool in_range(long long, long long, int) { return true; }
  static __device__ __forceinline__ double to_lds_wide(__int128 d, int shift) {
    return i128_to_double(d) * ldexp(1.0, -shift);
  }
  static __device__ __forceinline__ bool in_range_wide(__int128 d, int shift) {
    const __int128 q = d >> shift;
    return q > -1099511627776LL && q < 1099511627776LL;
  }
};

}  // namespace stormm_amd
