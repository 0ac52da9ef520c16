// Split fixed-precision numbers of 95 bits: a 64-bit word x and a 32-bit overflow word y with
// value = y * 2^63 + x, the representation of the reference's coordinates above 38 bits and
// forces above 40 bits (src/Synthesis/phasespace_synthesis.h:162,174; conversion and carry rules
// of src/Numerics/accumulation.cui:121-147,266-403).  The value of a pair is unique but its split
// is not: any y with x = value - y 2^63 representable is valid, and every routine here accepts
// any split.  Arithmetic on values goes through 128-bit integers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace stormm_amd {

constexpr double kTwo63 = 9223372036854775808.0;

__host__ __device__ inline __int128 i95_value(long long x, int y) {
  return (static_cast<__int128>(y) << 63) + static_cast<__int128>(x);
}

// A value (|v| < 2^94) split with x in [0, 2^63).
__host__ __device__ inline void i95_split(__int128 v, long long *x, int *y) {
  const __int128 hi = v >> 63;
  *y = static_cast<int>(hi);
  *x = static_cast<long long>(v - (hi << 63));
}

// A real number already on the fixed-precision grid, rounded to the nearest integer and split as
// the reference's doubleToInt95 does (the overflow word takes the whole multiples of 2^63).
__host__ __device__ inline void i95_from_double(double d, long long *x, int *y) {
  if (fabs(d) >= kTwo63) {
    const int spill = static_cast<int>(d / kTwo63);
    *x = static_cast<long long>(__builtin_rint(d - static_cast<double>(spill) * kTwo63));
    *y = spill;
  }
  else {
    *x = static_cast<long long>(__builtin_rint(d));
    *y = 0;
  }
}

// The real value of a pair, y 2^63 + x in double precision (int95ToDouble).
__host__ __device__ inline double i95_to_double(long long x, int y) {
  return (static_cast<double>(y) * kTwo63) + static_cast<double>(x);
}

// A 128-bit integer (|v| < 2^94) as the nearest-rounded double of its upper part plus its lower
// 32 bits: exact below 2^53, within one unit of the last place above.
__host__ __device__ inline double i128_to_double(__int128 v) {
  const long long hi = static_cast<long long>(v >> 32);
  const unsigned lo = static_cast<unsigned>(static_cast<unsigned long long>(v) & 0xffffffffULL);
  return (static_cast<double>(hi) * 4294967296.0) + static_cast<double>(lo);
}

// a + b with the carry of the 64-bit word moved into the overflow word (splitFPSum).
__host__ __device__ inline void i95_add(long long ax, int ay, long long bx, int by, long long *rx,
                                        int *ry) {
  const long long sx = static_cast<long long>(static_cast<unsigned long long>(ax) +
                                              static_cast<unsigned long long>(bx));
  int sy = ay + by;
  const bool carry = ((ax ^ sx) < 0) && ((ax ^ bx) >= 0);
  sy += carry ? ((bx < 0) ? -2 : 2) : 0;
  *rx = sx;
  *ry = sy;
}

}  // namespace stormm_amd
