// Work-unit evaluation stages (device functions).  A work unit is one thread block holding one
// or more whole systems in LDS.  The stages run in order, separated by block barriers:
//   stage_load      global coordinates / parameters -> LDS, accumulators zeroed
//   stage_psi       Born-radius integrals over 16 x 16 pair tiles (both directions per pair)
//   stage_radii     effective Born radii, GB self energies, self part of dE/dR (per atom, FP64)
//   stage_pairs     Coulomb + Lennard-Jones (exclusion masked) and GB pair energy / force / dE/dR
//   stage_chain     OBC chain rule on dE/dR (per atom, FP64)
//   stage_deriv     forces from the radius derivatives over the pair tiles
//   stage_valence   bonds, angles, cosine dihedrals and attenuated 1:4 pairs
// Forces, psi and dE/dR are accumulated in LDS as 64-bit fixed-precision integers, so the
// result does not depend on which lanes or work units processed which tiles.
#pragma once

#include <type_traits>

#include "mm_device.h"

namespace stormm_amd {

template <typename T> struct StageCtx {
  WuLds<T> L;
  WuDesc wu;
  int tid, nthr;
  int crd_shift;     // right shift from global fixed precision to LDS coordinates
  int org_shift;     // right shift from global fixed precision to the stored origins
  T crd_scale;       // LDS coordinate unit in Angstroms (single precision)
  int cycle;         // minimization cycle of this evaluation (clash-damping schedule)
  int rst_step;      // restraint step number when the kit gives no per-system cycle counts
  bool gb;
  bool move_pending = false;  // the coordinate stage adds L.move[slot] x direction first
  __device__ StageCtx(unsigned char *smem, int maxa, int waves) : L(smem, maxa, waves) {}
};

// Slot s of the unit's system that holds item t, given ascending per-slot starts pre[0..n]
// (pre[0] = 0 <= t < pre[n]): the s with pre[s] <= t < pre[s + 1].
__device__ __forceinline__ int find_slot(const int *pre, int n, int t) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= t) {
      lo = mid;
    }
    else {
      hi = mid;
    }
  }
  return lo;
}

// ------------------------------------------------------------------------------------------
// Work-unit set-up: system table, per-atom parameters, and the coordinate origins.
template <typename T>
__device__ __forceinline__ void stage_load_params(StageCtx<T> &c, const MMKit &k) {
  WuLds<T> &L = c.L;
  const WuDesc &wu = c.wu;
  if (c.tid < wu.sys_count) {
    L.sys[c.tid] = k.wu_sys[wu.sys_first + c.tid];
  }
  __syncthreads();
  if (c.tid == 0) {
    // Valence terms are taken kind by kind (all bonds of the unit, then all angles, then all
    // dihedrals), so that a chunk of 64 runs one kind's code.
    constexpr int K1 = kMaxWuSystems + 1;
    int cl = 0, o = 0, d = 0, vb = 0, va = 0, vd = 0;
    for (int s = 0; s < wu.sys_count; s++) {
      const WuSys &ws = L.sys[s];
      L.cpre[s] = cl;
      L.opre[s] = o;
      L.dpre[s] = d;
      L.vpre[s] = vb;
      L.vpre[K1 + s] = va;
      L.vpre[2 * K1 + s] = vd;
      cl += ws.clean_count;
      o += ws.offd_count - ws.clean_count;
      d += ws.diag_count;
      vb += ws.bond_count;
      va += ws.angl_count;
      vd += ws.term_count - ws.bond_count - ws.angl_count;
    }
    L.cpre[wu.sys_count] = cl;
    L.opre[wu.sys_count] = o;
    L.dpre[wu.sys_count] = d;
    L.vpre[wu.sys_count] = vb;
    L.vpre[K1 + wu.sys_count] = va;
    L.vpre[2 * K1 + wu.sys_count] = vd;
  }
  __syncthreads();
  // The unit's tile tasks, read once per evaluation: the clean off-diagonal tasks of every
  // system, then their other off-diagonal tasks, then the diagonal ones, each with its system's
  // LDS base added and its slot in the low bits of y.
  const int nclean = L.cpre[wu.sys_count];
  for (int t = c.tid; t < wu.task_count; t += c.nthr) {
    int s, idx;
    if (t < nclean) {
      s = find_slot(L.cpre, wu.sys_count, t);
      idx = L.sys[s].task_first + (t - L.cpre[s]);
    }
    else if (t < wu.offd_count) {
      const int to = t - nclean;
      s = find_slot(L.opre, wu.sys_count, to);
      idx = L.sys[s].task_first + L.sys[s].clean_count + (to - L.opre[s]);
    }
    else {
      const int td = t - wu.offd_count;
      s = find_slot(L.dpre, wu.sys_count, td);
      idx = L.sys[s].task_first + L.sys[s].offd_count + (td - L.dpre[s]);
    }
    int4 tk = k.tasks[idx];
    const int base = L.sys[s].lds_base;
    tk.x += base | (base << 16);
    tk.y |= s;
    L.tasks[t] = tk;
  }
  for (int a = c.tid; a < wu.lds_atoms; a += c.nthr) {
    int s = 0;
    for (int t = 1; t < wu.sys_count; t++) {
      s = (a >= L.sys[t].lds_base) ? t : s;
    }
    const WuSys ws = L.sys[s];
    const int local = a - ws.lds_base;
    const bool valid = local < ws.natom;
    const int g = ws.global_start + local;
    L.gidx[a] = valid ? g : -1;
    L.slot[a] = s;
    if (valid) {
      L.q[a] = static_cast<T>(k.q[g]);
      if (c.wu.lj_table) {
        L.p1[a] = static_cast<T>(k.lj_type[g]);
        L.p2[a] = T(0);
      }
      else {
        L.p1[a] = static_cast<T>(0.5 * k.ljp1[g]);  // half sigma: sigma_ij = p1_i + p1_j
        L.p2[a] = static_cast<T>(k.ljp2[g]);
      }
      const double rho = k.rho[g];
      L.rho[a] = static_cast<T>(rho);
      L.irho[a] = (rho > 0.0) ? static_cast<T>(1.0 / rho) : T(0);
      L.scr[a] = static_cast<T>(k.scr[g]);
    }
    else {
      L.q[a] = T(0);
      L.p1[a] = T(0);
      L.p2[a] = T(0);
      L.rho[a] = T(1);
      L.irho[a] = T(1);
      L.scr[a] = T(0);
    }
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ bool clash_params(const MMKit &k, int sys_index, int cycle, T &d,
                                             T &r);

// Coordinates into LDS (relative to a per-system origin) and all accumulators zeroed.
template <typename T, bool CLASH>
__device__ void stage_load_coords(StageCtx<T> &c, const MMKit &k) {
  WuLds<T> &L = c.L;
  const WuDesc &wu = c.wu;
  if (c.tid < wu.sys_count) {
    if (CLASH) {
      T d, r;
      const bool on = clash_params<T>(k, L.sys[c.tid].sys_index, c.cycle, d, r);
      L.clash_d[c.tid] = on ? d : T(0);
      L.clash_r[c.tid] = on ? r : T(0);
      L.clash_on[c.tid] = on ? 1 : 0;
    }
    const int gs = L.sys[c.tid].global_start;
    if (k.xcrd_ovrf != nullptr) {
      // 95-bit coordinates: the first atom's value with the bits below the LDS resolution
      // cleared, kept in units of 2^-(gpos_bits - org_shift) A.
      const long long *crd[3] = {k.xcrd, k.ycrd, k.zcrd};
      const int *ovr[3] = {k.xcrd_ovrf, k.ycrd_ovrf, k.zcrd_ovrf};
      for (int d = 0; d < 3; d++) {
        const __int128 v = (i95_value(crd[d][gs], ovr[d][gs]) >> c.crd_shift) << c.crd_shift;
        L.org[3 * c.tid + d] = static_cast<long long>(v >> c.org_shift);
      }
    }
    else {
      const long long mask = ~((1LL << c.crd_shift) - 1LL);
      long long x0 = k.xcrd[gs], y0 = k.ycrd[gs], z0 = k.zcrd[gs];
      if (c.move_pending) {
        // The origin comes from the moved first atom, as it would after a separate move.  This
        // thread wrote the system's move factor itself before the stage began.
        const double mf = L.move[c.tid];
        if (mf != 0.0) {
          x0 += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.xfrc[gs])));
          y0 += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.yfrc[gs])));
          z0 += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.zfrc[gs])));
        }
      }
      L.org[3 * c.tid + 0] = x0 & mask;
      L.org[3 * c.tid + 1] = y0 & mask;
      L.org[3 * c.tid + 2] = z0 & mask;
    }
  }
  for (int e = c.tid; e < wu.sys_count * 16; e += c.nthr) {
    L.nrg[e] = 0;
  }
  if (c.tid < kStageCounters) {
    L.work[c.tid] = 0;
  }
  __syncthreads();
  for (int a = c.tid; a < wu.lds_atoms; a += c.nthr) {
    const int g = L.gidx[a];
    const int s = L.slot[a];
    L.psi[a] = 0;
    L.dedr[a] = 0;
    L.fx[a] = 0;
    L.fy[a] = 0;
    L.fz[a] = 0;
    L.rad[a] = T(1);
    L.irad[a] = T(1);
    L.sd[a] = T(0);
    if (g >= 0 && k.xcrd_ovrf != nullptr) {
      const __int128 dx = i95_value(k.xcrd[g], k.xcrd_ovrf[g]) -
                          (static_cast<__int128>(L.org[3 * s]) << c.org_shift);
      const __int128 dy = i95_value(k.ycrd[g], k.ycrd_ovrf[g]) -
                          (static_cast<__int128>(L.org[3 * s + 1]) << c.org_shift);
      const __int128 dz = i95_value(k.zcrd[g], k.zcrd_ovrf[g]) -
                          (static_cast<__int128>(L.org[3 * s + 2]) << c.org_shift);
      if (!(CrdConv<T>::in_range_wide(dx, c.crd_shift) &&
            CrdConv<T>::in_range_wide(dy, c.crd_shift) &&
            CrdConv<T>::in_range_wide(dz, c.crd_shift))) {
        atomicOr(k.status, static_cast<int>(kStatusCoordinateRange));
      }
      L.cx[a] = CrdConv<T>::to_lds_wide(dx, c.crd_shift);
      L.cy[a] = CrdConv<T>::to_lds_wide(dy, c.crd_shift);
      L.cz[a] = CrdConv<T>::to_lds_wide(dz, c.crd_shift);
    }
    else if (g >= 0) {
      long long gx = k.xcrd[g], gy = k.ycrd[g], gz = k.zcrd[g];
      if (c.move_pending) {
        // The pending line move, x += rint(mf * direction) with the normalized direction in the
        // force arrays (line_movement.cui:343-360), written back as the new coordinates.
        const double mf = L.move[s];
        if (mf != 0.0) {
          gx += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.xfrc[g])));
          gy += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.yfrc[g])));
          gz += static_cast<long long>(__builtin_rint(mf * static_cast<double>(k.zfrc[g])));
          k.xcrd[g] = gx;
          k.ycrd[g] = gy;
          k.zcrd[g] = gz;
        }
      }
      const long long ox = L.org[3 * s], oy = L.org[3 * s + 1], oz = L.org[3 * s + 2];
      if (!(CrdConv<T>::in_range(gx, ox, c.crd_shift) &&
            CrdConv<T>::in_range(gy, oy, c.crd_shift) &&
            CrdConv<T>::in_range(gz, oz, c.crd_shift))) {
        atomicOr(k.status, static_cast<int>(kStatusCoordinateRange));
      }
      L.cx[a] = CrdConv<T>::to_lds(gx, ox, c.crd_shift);
      L.cy[a] = CrdConv<T>::to_lds(gy, oy, c.crd_shift);
      L.cz[a] = CrdConv<T>::to_lds(gz, oz, c.crd_shift);
    }
    else {
      L.cx[a] = 0;
      L.cy[a] = 0;
      L.cz[a] = 0;
    }
  }
  __syncthreads();
}

template <typename T>
__device__ void init_ctx(StageCtx<T> &c, const MMKit &k, int wu_index, int terms) {
  c.wu = k.wu[wu_index];
  c.tid = threadIdx.x;
  c.nthr = blockDim.x;
  if (sizeof(T) == 4) {
    c.crd_shift = (k.gpos_bits > kLdsCrdBits) ? (k.gpos_bits - kLdsCrdBits) : 0;
    c.crd_scale = T(ldexp(1.0, -(k.gpos_bits - c.crd_shift)));
  }
  else {
    c.crd_shift = k.gpos_bits;
    c.crd_scale = T(1);
  }
  c.org_shift = (k.xcrd_ovrf != nullptr) ? (k.gpos_bits - kOriginBits) : 0;
  c.gb = (k.igb != 0) && ((terms & kTermNonbonded) != 0);
  c.cycle = k.clash_cycle;
  c.rst_step = k.rst_step;
}

// Clash-damping parameters of one system for this evaluation
// (src/MolecularMechanics/hpc_minimization.cu:269-274): the floor and ratio shrink linearly to
// zero over the first clash_cycles cycles.  The damped potentials apply when either value
// exceeds 10^-6, the threshold at which the reference launches its damped kernels
// (src/Potential/hpc_nonbonded_potential.cu:1476 in double precision, :1662 in single).
template <typename T>
__device__ __forceinline__ bool clash_params(const MMKit &k, int sys_index, int cycle,
                                             T &dist, T &ratio) {
  double d = k.clash_r0, r = k.clash_ratio;
  if (k.clash_cycles > 0) {
    const int i = (k.clash_step != nullptr) ? k.clash_step[sys_index] + k.clash_step_offset
                                            : cycle;
    if (i < k.clash_cycles) {
      const double progress = static_cast<double>(k.clash_cycles - i) /
                              static_cast<double>(k.clash_cycles);
      d *= progress;
      r *= progress;
    }
    else {
      d = 0.0;
      r = 0.0;
    }
  }
  dist = static_cast<T>(d);
  ratio = static_cast<T>(r);
  if (sizeof(T) == 8) {
    return d >= 1.0e-6 || r >= 1.0e-6;
  }
  return static_cast<float>(d) > 0.000001f || static_cast<float>(r) > 0.000001f;
}

// Rotation-step mask of a 16-bit mask over the j atoms of a tile: bit ks of the result is bit
// (il - ks) mod 16 of jmask, the j atom lane il holds after ks rotations.  The reversed mask,
// rotated right by 15 - il, puts those bits in step order.
__device__ __forceinline__ unsigned step_mask(unsigned jmask, int il) {
  const unsigned r = __brev(jmask) >> 16;
  const int sh = (15 - il) & 15;
  return ((r >> sh) | (r << (16 - sh))) & 0xffffu;
}

// Number of rotation steps the wave runs for its current tasks: 16 when any of its rows holds
// an off-diagonal tile, 8 when all hold diagonal tiles.
__device__ __forceinline__ int wave_steps(bool diag) {
  const unsigned long long offd = __ballot(!diag);
  return (offd != 0ULL) ? 16 : 8;
}

// A tile task as seen by one lane of its 16-lane row.  Lane il starts with j lane jl0: il on an
// off-diagonal tile, il - 1 (mod 16) on a diagonal one, so that a diagonal tile's step ks meets
// the atom ks + 1 lanes back and its first 8 steps hold every pair of the tile once (the 8th
// step, which meets each pair twice, on lanes 0..7 only) and no atom paired with itself.  Bit ks
// of `valid` is set when the lane's pair at rotation step ks exists.
struct TileTask {
  int ib, jb, slot, ni, nj, i, jl0, j0, nsteps;
  bool diag, ivalid;
  unsigned valid;
  __device__ __forceinline__ TileTask(const Task &tk, int il) {
    ib = tk.x & 0xffff;
    jb = tk.x >> 16;
    slot = tk.y & 0xff;
    diag = ((tk.y >> 8) & 1) != 0;
    ni = (tk.y >> 12) & 0xff;
    nj = (tk.y >> 20) & 0xff;
    ivalid = il < ni;
    i = ib + il;
    jl0 = (il - (diag ? 1 : 0)) & 15;
    j0 = jb + jl0;
    const unsigned jm = (nj >= 16) ? 0xffffu : ((1u << nj) - 1u);
    unsigned v = step_mask(jm, jl0);
    if (diag) {
      v &= (il < 8) ? 0xffu : 0x7fu;
    }
    valid = ivalid ? v : 0u;
    nsteps = wave_steps(diag);
  }
  // The j lane whose sums a lane holds after the last step (the j sums move one lane per step).
  __device__ __forceinline__ int jl_home() const {
    return (jl0 - (nsteps - 1)) & 15;
  }
};

// Coordinates of lane il's i atom and first j atom of a tile, in Angstroms.  The padding slots
// of a partial tile (lanes past ni on the i side, past nj on the j side) are put far away, so
// that their pairs are far-field pairs with terms of exactly zero in the Born-integral and
// derivative loops, which need no pair mask on off-diagonal tiles.
template <typename T>
__device__ __forceinline__ void tile_coords(const StageCtx<T> &c, const TileTask &tt, T &xi,
                                            T &yi, T &zi, T &xj, T &yj, T &zj) {
  const WuLds<T> &L = c.L;
  const auto ox = tile_centre(L.cx[tt.ib], L.cx[tt.jb]);
  const auto oy = tile_centre(L.cy[tt.ib], L.cy[tt.jb]);
  const auto oz = tile_centre(L.cz[tt.ib], L.cz[tt.jb]);
  const bool jvalid = tt.jl0 < tt.nj;
  xi = tt.ivalid ? tile_crd(L.cx[tt.i], ox, c.crd_scale) : FarAway<T>::kCrd;
  yi = tt.ivalid ? tile_crd(L.cy[tt.i], oy, c.crd_scale) : FarAway<T>::kCrd;
  zi = tt.ivalid ? tile_crd(L.cz[tt.i], oz, c.crd_scale) : FarAway<T>::kCrd;
  xj = jvalid ? tile_crd(L.cx[tt.j0], ox, c.crd_scale) : -FarAway<T>::kCrd;
  yj = jvalid ? tile_crd(L.cy[tt.j0], oy, c.crd_scale) : -FarAway<T>::kCrd;
  zj = jvalid ? tile_crd(L.cz[tt.j0], oz, c.crd_scale) : -FarAway<T>::kCrd;
}

// True on every lane when `b` holds on every active lane of the wavefront.
__device__ __forceinline__ bool wave_all(bool b) {
  return __ballot(!b) == 0ULL;
}

__device__ __forceinline__ bool step_bit(unsigned mask, int ks) {
  return ((mask >> ks) & 1u) != 0u;
}

// ------------------------------------------------------------------------------------------

// Tile task t of the work unit, placed at its system's LDS base and tagged with its slot
// (staged in LDS by stage_load_params).
template <typename T>
__device__ __forceinline__ Task fetch_task(const StageCtx<T> &c, const MMKit &, int t) {
  return c.L.tasks[t];
}

// Task dealing for the Born-integral stages: pass tb gives the four 16-lane rows of a
// wavefront four consecutive tasks; rows past the end of the list get an empty diagonal task,
// so every lane of a wavefront runs the same passes and rotation steps (the near queue below
// needs the whole wavefront at each step).
template <typename T>
__device__ __forceinline__ Task wave_task(const StageCtx<T> &c, const MMKit &k, int t) {
  if (t < c.wu.task_count) {
    return fetch_task(c, k, t);
  }
  return make_int4(0, 1 << 8, -1, 0);
}

// Work items of a stage are taken by the wavefronts of the block one at a time from a counter in
// LDS, so a wavefront that finishes early takes more items (the result does not depend on which
// wavefront took which item: every contribution goes to integer accumulators).
__device__ __forceinline__ int next_item(int *counter, int lane) {
  int item = 0;
  if (lane == 0) {
    item = atomicAdd(counter, 1);
  }
  return __builtin_amdgcn_readfirstlane(item);
}

// Item i of a stage holding `quads` tile-task quads and `chunks` chunks of 64 valence terms, the
// chunks spread evenly among the quads so that latency-bound valence work overlaps the
// arithmetic of the tiles on the other wavefronts.  Returns the chunk index for a valence item
// and -1 - (quad index) for a tile item.
__device__ __forceinline__ int stage_item(int i, int quads, int chunks) {
  if (chunks == 0) {
    return -1 - i;
  }
  const unsigned total = static_cast<unsigned>(quads + chunks);
  const int c0 = static_cast<int>((static_cast<unsigned>(i) * chunks) / total);
  const int c1 = static_cast<int>((static_cast<unsigned>(i + 1) * chunks) / total);
  return (c1 > c0) ? c0 : (-1 - (i - c0));
}

template <typename T, bool FORCES, bool CLASH>
__device__ void valence_term(StageCtx<T> &c, const MMKit &k, int kind, int t);

// Chunks of 64 valence terms of the unit: its bonds, then its angles, then its dihedrals, each
// kind starting a chunk of its own.
template <typename T> __device__ __forceinline__ int valence_chunks(const StageCtx<T> &c) {
  constexpr int K1 = kMaxWuSystems + 1;
  const int n = c.wu.sys_count;
  return ((c.L.vpre[n] + 63) >> 6) + ((c.L.vpre[K1 + n] + 63) >> 6) +
         ((c.L.vpre[2 * K1 + n] + 63) >> 6);
}

// Valence chunk `chunk` (wavefront-uniform), one term per lane.
template <typename T, bool FORCES, bool CLASH>
__device__ __forceinline__ void valence_chunk(StageCtx<T> &c, const MMKit &k, int chunk,
                                              int lane) {
  constexpr int K1 = kMaxWuSystems + 1;
  const int n = c.wu.sys_count;
  const int cb = (c.L.vpre[n] + 63) >> 6;
  const int ca = (c.L.vpre[K1 + n] + 63) >> 6;
  const int kind = (chunk < cb) ? kVBond : (chunk < cb + ca) ? kVAngle : kVDihedral;
  const int local = chunk - ((kind == kVBond) ? 0 : (kind == kVAngle) ? cb : cb + ca);
  valence_term<T, FORCES, CLASH>(c, k, kind, (local << 6) + lane);
}

// Near-field pair queue of a wavefront.  Entries pack the LDS indices of the pair (9 bits each)
// and which of the two directions (onto i, onto j) lies within 4 screened radii.
constexpr unsigned kNearI = 1u << 18;
constexpr unsigned kNearJ = 1u << 19;

__device__ __forceinline__ unsigned near_entry(int i, int j, bool near_i, bool near_j) {
  return static_cast<unsigned>(i) | (static_cast<unsigned>(j) << 9) | (near_i ? kNearI : 0u) |
         (near_j ? kNearJ : 0u);
}

// Appends the entries of the lanes with `need` set, in lane order; n is wave-uniform.  Every
// lane stores: the lanes with an entry to n, n + 1, ..., the others to the slots after them
// (the r-th lane without an entry, counting from lane 0, to n + 63 - r, which is at least
// n + count because at most 64 - count lanes have none), which later pushes overwrite (before a
// group of steps the queue holds at most 63 entries and a group pushes at most 4 x 64, so every
// store stays inside the queue's stride).  The step thus has no branch, which would split it
// and keep the rotated operands from folding into their instructions.  rlane4 is 4 (63 - lane),
// a byte offset.
__device__ __forceinline__ void nearq_push(unsigned *q, int &n, bool near_i, bool near_j,
                                           unsigned entry, int rlane4) {
  // The mask is the union of the two directions' ballots (each a comparison's lane mask).
  const bool need = near_i || near_j;
  const unsigned long long m = __builtin_amdgcn_ballot_w64(near_i) |
                               __builtin_amdgcn_ballot_w64(near_j);
  const int below = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(m >> 32),
                                              __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(m),
                                                                        0u));
  // For a lane without an entry, below + 63 - lane = 63 - r.
  unsigned char *slot = reinterpret_cast<unsigned char *>(q + n) + (below << 2);
  *reinterpret_cast<unsigned *>(slot + (need ? 0 : rlane4)) = entry;
  n += __popcll(m);
  __builtin_amdgcn_wave_barrier();
}

// The Born-integral and derivative stages run their rotation steps in groups of kStepGroup with
// compile-time rotation counts inside a group; the j-side data move kStepGroup lanes between
// groups, and the near-pair queue is drained after each group, taking 64 entries at a time from
// its end.  Before a group the queue holds at most 63 entries, so a group's last push starts at
// entry 63 + (kStepGroup - 1) x 64 at most and stores its 64 lanes below that plus 64.
constexpr int kStepGroup = 4;
static_assert(63 + ((kStepGroup - 1) * 64) + 64 <= kNearQueue,
              "near queue too short for a step group");

// Geometry of a queued pair, recomputed from the LDS coordinates.
template <typename T> struct NearPair {
  int i, j;
  bool near_i, near_j;
  T dx, dy, dz, r2, invr, r;
  __device__ __forceinline__ NearPair(const StageCtx<T> &c, unsigned e) {
    using P = Prec<T>;
    i = static_cast<int>(e & 0x1ffu);
    j = static_cast<int>((e >> 9) & 0x1ffu);
    near_i = (e & kNearI) != 0u;
    near_j = (e & kNearJ) != 0u;
    dx = crd_delta(c.L.cx[j], c.L.cx[i], c.crd_scale);
    dy = crd_delta(c.L.cy[j], c.L.cy[i], c.crd_scale);
    dz = crd_delta(c.L.cz[j], c.L.cz[i], c.crd_scale);
    r2 = (dx * dx) + (dy * dy) + (dz * dz);
    invr = P::rsqrt(r2);
    r = r2 * invr;
  }
};

// Near-field Born-integral terms of the first `count` queued pairs, one pair per lane.
template <typename T>
__device__ __forceinline__ void psi_drain(StageCtx<T> &c, const unsigned *q, int count,
                                          int lane) {
  WuLds<T> &L = c.L;
  if (lane < count) {
    const NearPair<T> p(c, q[lane]);
    if (p.near_i) {
      const T t = psi_near(p.r, p.r2, p.invr, L.rho[p.i], L.irho[p.i], L.scr[p.j]);
      lds_add(&L.psi[p.i], fixed_contrib(t, kPsiBits));
    }
    if (p.near_j) {
      const T t = psi_near(p.r, p.r2, p.invr, L.rho[p.j], L.irho[p.j], L.scr[p.i]);
      lds_add(&L.psi[p.j], fixed_contrib(t, kPsiBits));
    }
  }
}

// Born-radius integrals (src/Potential/nonbonded_potential.tpp:560-650).  Each lane evaluates
// the far-field series of its pair at every rotation step and queues the pair when either
// direction is within 4 screened radii; full wavefronts of queued pairs are evaluated together.
// The series terms of one task are summed in the working precision (at most 16 small terms per
// lane) and added to the fixed-precision accumulators once per task.  With x = (s / r)^2 for the
// sphere of screened radius s on the other atom, the far-field term is -s^3 / r^4 times the
// series in x, and r > 4 s is x < 1/16.
// Chunks of valence terms (`val_chunks`, possibly 0) are taken as work items alongside the tiles.
template <typename T, bool FORCES, bool CLASH>
__device__ void stage_psi(StageCtx<T> &c, const MMKit &k, int val_chunks) {
  using P = Prec<T>;
  WuLds<T> &L = c.L;
  const int il = c.tid & 15;
  const int lane = c.tid & 63;
  const int rlane4 = (63 - lane) << 2;
  // Wavefront-uniform, so the queue's base address stays in scalar registers.
  const int wave = __builtin_amdgcn_readfirstlane(c.tid >> 6);
  unsigned *q = L.nearq + (wave * kNearQueueStride);
  int qn = 0;
  const double psi_scale = ldexp(1.0, kPsiBits);
  const int quads = (c.wu.task_count + 3) >> 2;
  const int items = quads + val_chunks;
  for (int item = next_item(&L.work[0], lane); item < items; item = next_item(&L.work[0], lane)) {
    const int it = stage_item(item, quads, val_chunks);
    if (it >= 0) {
      valence_chunk<T, FORCES, CLASH>(c, k, it, lane);
      continue;
    }
    const TileTask tt(wave_task(c, k, ((-1 - it) << 2) + ((c.tid >> 4) & 3)), il);
    T xi, yi, zi, xj, yj, zj;
    tile_coords(c, tt, xi, yi, zi, xj, yj, zj);
    const T si = L.scr[tt.i];
    const T si2 = si * si;
    const T si3 = si2 * si;
    const T sj = L.scr[tt.j0];
    T sj2 = sj * sj;
    T sj3 = sj2 * sj;
    // Queue entry bits of the pair: i in bits 0..8, the j atom in bits 9..17.
    const unsigned ei = static_cast<unsigned>(tt.i);
    int ej = tt.j0 << 9;
    T acci = T(0), accj = T(0);
    unsigned vbits = tt.valid;
    // Off-diagonal tiles need no pair mask (their padding atoms sit far away); a wavefront whose
    // four tiles are all off-diagonal runs the 16 steps without one.  On diagonal tiles the mask
    // removes each lane's own atom and the second copy of each pair.
    auto run = [&](auto mc) {
      constexpr bool MASKED = decltype(mc)::value;
      const int nsteps = MASKED ? tt.nsteps : kTile;
      for (int g = 0; g < nsteps; g += kStepGroup) {
        auto step = [&](auto kc) {
          constexpr int K = decltype(kc)::value;
          if (MASKED && g + K >= nsteps) {
            return;
          }
          // An absent pair is put far away: both directions are then far-field with terms
          // that round to exactly zero, so no validity test follows.
          const bool valid = !MASKED || (((vbits >> K) & 1u) != 0u);
          const T dx = row_ror<K>(xj) - xi;
          const T dy = row_ror<K>(yj) - yi;
          const T dz = row_ror<K>(zj) - zi;
          const T r2c = (dx * dx) + (dy * dy) + (dz * dz);
          const T r2 = valid ? r2c : FarAway<T>::kR2;
          const T invr2 = P::rcp(r2);
          const T invr4 = invr2 * invr2;
          const T x_i = row_ror<K>(sj2) * invr2;
          const T x_j = si2 * invr2;
          const bool far_i = x_i < T(0.0625);
          const bool far_j = x_j < T(0.0625);
          const T ti = psi_series(x_i) * (row_ror<K>(sj3) * invr4);
          const T tj = psi_series(x_j) * (si3 * invr4);
          acci -= far_i ? ti : T(0);
          accj = row_rotate(accj) - (far_j ? tj : T(0));
          nearq_push(q, qn, !far_i, !far_j,
                     ei | static_cast<unsigned>(row_ror<K>(ej)) | (far_i ? 0u : kNearI) |
                         (far_j ? 0u : kNearJ),
                     rlane4);
        };
        unrolled_steps(step, std::integer_sequence<int, 0, 1, 2, 3>{});
        settle(acci);
        vbits >>= kStepGroup;
        xj = row_ror<kStepGroup>(xj);
        yj = row_ror<kStepGroup>(yj);
        zj = row_ror<kStepGroup>(zj);
        sj2 = row_ror<kStepGroup>(sj2);
        sj3 = row_ror<kStepGroup>(sj3);
        ej = row_ror<kStepGroup>(ej);
        while (qn >= 64) {
          psi_drain(c, q + (qn - 64), 64, lane);
          qn -= 64;
        }
      }
    };
    if (wave_all(!tt.diag)) {
      run(std::false_type{});
    }
    else {
      run(std::true_type{});
    }
    if (tt.ivalid) {
      lds_add(&L.psi[tt.i], Prec<double>::fixed(static_cast<double>(acci) * psi_scale));
    }
    // After the last step the j sums sit nsteps - 1 lanes right of the lanes that loaded them.
    const int jl_home = tt.jl_home();
    if (jl_home < tt.nj) {
      lds_add(&L.psi[tt.jb + jl_home], Prec<double>::fixed(static_cast<double>(accj) *
                                                           psi_scale));
    }
  }
  psi_drain(c, q, qn, lane);
  __syncthreads();
}

// ------------------------------------------------------------------------------------------
// Effective radii (src/Potential/nonbonded_potential.tpp:731-781), self energies and the self
// part of the radius derivative accumulator (:784-806), all in double precision.
template <typename T, bool FORCES>
__device__ void stage_radii(StageCtx<T> &c, const MMKit &k, double coulomb_inv_scaled) {
  WuLds<T> &L = c.L;
  const double nrg_scale = ldexp(1.0, k.nrg_bits);
  const double inv_diel = 1.0 / k.dielectric;
  for (int a = c.tid; a < c.wu.lds_atoms; a += c.nthr) {
    const int g = L.gidx[a];
    if (g < 0) {
      continue;
    }
    const double psi = static_cast<double>(L.psi[a]) * ldexp(1.0, -kPsiBits);
    const double rho = k.rho[g];
    const double pbr = k.pbr[g];
    double R;
    if (k.igb == 1) {
      R = 1.0 / ((1.0 / rho) + psi);
      R = (R < 0.0) ? 30.0 : R;
    }
    else {
      const double fipsi = psi * (-rho);
      const double al = k.gb_alpha[g], be = k.gb_beta[g], ga = k.gb_gamma[g];
      const double thi = tanh((al - (be * fipsi) + (ga * fipsi * fipsi)) * fipsi);
      R = 1.0 / ((1.0 / rho) - thi / pbr);
      if (FORCES) {
        // dR/dpsi of the OBC model (nonbonded_potential.tpp:985-1017), kept for the chain rule
        // in the accumulator slot whose psi is no longer needed.
        const double dr_dpsi = (al - ((2.0 * be) - (3.0 * ga * fipsi)) * fipsi) *
                               (1.0 - thi * thi) * rho / pbr;
        L.psi[a] = __double_as_longlong(dr_dpsi);
      }
    }
    L.rad[a] = static_cast<T>(R);
    L.irad[a] = static_cast<T>(1.0 / R);
    // q holds charge * sqrt(Coulomb), so q^2 = Coulomb * charge^2.
    const double qg = k.q[g];
    const double expmkf = ((k.kappa != 0.0) ? exp(-kGbKscale * k.kappa * R) : 1.0) * inv_diel;
    const double dielfac = 1.0 - expmkf;
    const double atmq2h = 0.5 * qg * qg;
    const double atmqd2h = atmq2h * dielfac;
    const double contrib = -atmqd2h / R;
    lds_add(&L.nrg[L.slot[a] * 16 + kGb], static_cast<long long>(__builtin_rint(contrib *
                                                                                  nrg_scale)));
    if (FORCES) {
      const double sdi = atmqd2h - (kGbKscale * k.kappa * atmq2h * expmkf * R);
      L.dedr[a] = static_cast<long long>(__builtin_rint(sdi * ldexp(1.0, kDedrBits)));
    }
  }
  (void)coulomb_inv_scaled;
  __syncthreads();
}

// ------------------------------------------------------------------------------------------
// Coulomb + Lennard-Jones (src/Potential/nonbonded_potential.tpp:174-231,312-373) and the GB
// pair energy, force and radius derivative (:811-972).  Every lane evaluates its pair at every
// step; absent and excluded pairs are removed by zeroing the charge product and the LJ
// coefficients, and the squared distance of an absent pair is replaced by 1 to keep the
// arithmetic finite.  Forces are summed per lane in the working precision and converted to fixed
// precision once per task.
// Damped Coulomb and Lennard-Jones of one pair at distance r (clash damping,
// src/Potential/nonbonded_potential_tilegroups.cui:536-602): inside clash distance cd the
// Coulomb potential becomes the parabola qq (ca (r + 1)^2 + cb), and inside ratio * sigma the
// Lennard-Jones potential becomes a (r + 1)^4 + b, each meeting the plain potential and its
// slope at the boundary.  fmag receives dE/dr / r.
template <typename T, bool FORCES>
__device__ __forceinline__ void clash_pair(T r, T qq, T lja, T ljb, T sigma, T cd, T ca, T cb,
                                           T ratio, T &e_qq, T &e_lj, T &fmag) {
  const T tiny = T(1.0e-10);   // constants::tiny, src/Constants/scaling.h:26
  if (r < cd) {
    e_qq += qq * ((ca * (r + T(1)) * (r + T(1))) + cb);
    if (FORCES && r > tiny) {
      fmag += qq * ((T(2) * ca) + (T(2) * ca / r));
    }
  }
  else {
    e_qq += qq / r;
    if (FORCES) {
      fmag -= qq / (r * r * r);
    }
  }
  const T vdw_limit = ratio * sigma;
  if (r < vdw_limit) {
    const T invrlim = T(1) / vdw_limit;
    const T invrlim2 = invrlim * invrlim;
    const T invrlim6 = invrlim2 * invrlim2 * invrlim2;
    const T aparm = (((T(6) * ljb) - (T(12) * lja * invrlim6)) * invrlim * invrlim6) /
                    ((((((T(4) * vdw_limit) + T(12)) * vdw_limit) + T(12)) * vdw_limit) + T(4));
    const T r_plus_one = r + T(1);
    const T arpo_three = aparm * r_plus_one * r_plus_one * r_plus_one;
    const T rlimit_plus_one = vdw_limit + T(1);
    const T arlimit_po_four = aparm * (rlimit_plus_one * rlimit_plus_one * rlimit_plus_one *
                                       rlimit_plus_one);
    const T bparm = (((lja * invrlim6) - ljb) * invrlim6) - arlimit_po_four;
    e_lj += (arpo_three * r_plus_one) + bparm;
    if (FORCES && r > tiny) {
      fmag += T(4) * arpo_three / r;
    }
  }
  else {
    const T invr2 = T(1) / (r * r);
    const T invr4 = invr2 * invr2;
    e_lj += ((lja * invr4 * invr2) - ljb) * invr4 * invr2;
    if (FORCES) {
      fmag += ((T(6) * ljb) - (T(12) * lja * invr4 * invr2)) * invr4 * invr4;
    }
  }
}

// CLASH selects the kernels that damp Coulomb and Lennard-Jones for the systems whose clash
// flag is set in this evaluation; the other systems take the plain potentials.  The rotation
// steps run in groups of kStepGroup with compile-time rotation counts, as in stage_psi: within a
// group the j-side data stay put and are read through DPP operands; the j-side accumulators move
// one lane per step.  The radius-derivative sums hold dE/dR / R and are scaled
// by the atom's radius at the end.
template <typename T, bool FORCES, bool LJTABLE, bool GB, bool CLASH>
__device__ void stage_pairs(StageCtx<T> &c, const MMKit &k) {
  using P = Prec<T>;
  WuLds<T> &L = c.L;
  const int row = c.tid >> 4;
  const int nrows = c.nthr >> 4;
  const int il = c.tid & 15;
  const T frc_scale = T(ldexp(1.0, k.frc_bits));
  const T dedr_scale = T(ldexp(1.0, kDedrBits));
  const T nrg_scale = T(ldexp(1.0, k.nrg_bits));
  const T inv_diel = T(1.0 / k.dielectric);
  const T kap = T(k.kappa * kGbKscale);
  const bool use_kappa = (k.kappa != 0.0);
  const T dielfac0 = T(1) - inv_diel;
  for (int t = row; t < c.wu.task_count; t += nrows) {
    const Task tk = fetch_task(c, k, t);
    const TileTask tt(tk, il);
    const unsigned emask = (tk.z >= 0 && tt.ivalid) ? k.masks[tk.z + il] : 0u;
    const unsigned nbmask = tt.valid & ~step_mask(emask, tt.jl0);
    T xi, yi, zi, xj, yj, zj;
    tile_coords(c, tt, xi, yi, zi, xj, yj, zj);
    const T qi = L.q[tt.i], p1i = L.p1[tt.i], p2i = L.p2[tt.i], Ri = L.rad[tt.i];
    T qj = L.q[tt.j0], p1j = L.p1[tt.j0], p2j = L.p2[tt.j0], Rj = L.rad[tt.j0];
    // exp(-r^2 / (4 Ri Rj)) = GbExp(r^2 * (1 / Rj) * (scale / Ri)).
    T iRj = GB ? L.irad[tt.j0] : T(0);
    const T iRi = GB ? L.irad[tt.i] * GbExp<T>::kScale : T(0);
    T fxi = T(0), fyi = T(0), fzi = T(0), dri = T(0);
    T fxj = T(0), fyj = T(0), fzj = T(0), drj = T(0);
    // Each lane sums its energies over the task in the calculation precision and converts the
    // sums to fixed precision once, as the reference does per tile
    // (src/Potential/nonbonded_potential_tilegroups.cui:461-464,640-645).
    T e_qq = T(0), e_lj = T(0), e_gb = T(0);
    // The row of the system's A / B table that atom i's type selects (a per-lane pointer, so
    // the steps need no kernel-argument registers to address it).
    const double2 *ljrow = nullptr;
    if (LJTABLE) {
      ljrow = k.ljab + L.sys[tt.slot].lj_off +
              (static_cast<int>(p1i) * L.sys[tt.slot].lj_nt);
    }
    bool clash_on = false;
    T cd = T(0), ca = T(0), cb = T(0), cratio = T(0);
    if (CLASH) {
      clash_on = (L.clash_on[tt.slot] != 0);
      cd = L.clash_d[tt.slot];
      cratio = L.clash_r[tt.slot];
      ca = T(-0.5) / (cd * cd * (cd + T(1)));
      cb = (T(1) / cd) - (ca * (cd + T(1)) * (cd + T(1)));
    }
    // The clash-damped kernels evaluate both potentials in one step (damping is per system, so a
    // wavefront's rows can differ): they rotate every step to stay within the register budget.
    constexpr int kGroup = CLASH ? 1 : kStepGroup;
    unsigned vbits = tt.valid, nbbits = nbmask;
    // A wavefront whose four tiles are all clean (off-diagonal, 16 x 16 atoms, no excluded pair)
    // runs the 16 steps without pair masks.
    auto run = [&](auto mc, auto kc) {
      constexpr bool MASKED = decltype(mc)::value;
      constexpr bool KAPPA = decltype(kc)::value;
      const int nsteps = MASKED ? tt.nsteps : kTile;
      for (int g = 0; g < nsteps; g += kGroup) {
        auto step = [&](auto sc) {
          constexpr int K = decltype(sc)::value;
          if (MASKED && g + K >= nsteps) {
            return;
          }
          const bool valid = !MASKED || (((vbits >> K) & 1u) != 0u);
          const bool nbok = !MASKED || (((nbbits >> K) & 1u) != 0u);
          const T dx = row_ror<K>(xj) - xi;
          const T dy = row_ror<K>(yj) - yi;
          const T dz = row_ror<K>(zj) - zi;
          const T r2c = (dx * dx) + (dy * dy) + (dz * dz);
          const T r2 = valid ? r2c : T(1);
          const T qiqj = row_ror<K>(qj) * qi;
          T fmag = T(0);
          if (GB) {
            const T qgb = valid ? qiqj : T(0);
            const T rr = row_ror<K>(Rj) * Ri;
            const T efac = GbExp<T>::eval(r2 * (row_ror<K>(iRj) * iRi));
            const T fgb2 = r2 + (rr * efac);
            const T fgbi = P::rsqrt(fgb2);
            // egb = qq (1 - exp(-kappa fgb) / diel) / fgb, temp6 = -(dE/dfgb) / fgb.
            T egb, temp6;
            if (KAPPA) {
              const T fgbk = -kap * fgb2 * fgbi;
              const T expmkf = P::exp(fgbk) * inv_diel;
              const T dielfac = T(1) - expmkf;
              const T qf = qgb * fgbi;
              egb = qf * dielfac;
              temp6 = qf * fgbi * fgbi * (dielfac + (fgbk * expmkf));
            }
            else {
              egb = qgb * fgbi * dielfac0;
              temp6 = egb * fgbi * fgbi;
            }
            e_gb -= egb;
            if (FORCES) {
              fmag += temp6 * (T(1) - (T(0.25) * efac));
              const T temp5 = T(0.5) * efac * temp6 * (rr + (T(0.25) * r2));
              dri += temp5;
              drj = row_rotate(drj) + temp5;
            }
          }
          const T qnb = nbok ? qiqj : T(0);
          if (CLASH && clash_on) {
            const T pj1 = row_ror<K>(p1j);
            const T pj2 = row_ror<K>(p2j);
            T lja, ljb, sigma;
            if (LJTABLE) {
              const double2 ab = ljrow[static_cast<int>(pj1)];
              lja = nbok ? static_cast<T>(ab.x) : T(0);
              ljb = nbok ? static_cast<T>(ab.y) : T(0);
              sigma = (ljb > T(1.0e-6)) ? P::sqrt(P::cbrt(lja / ljb)) : T(0);
            }
            else {
              const T sig = p1i + pj1;
              const T sig2 = sig * sig;
              const T sig6 = sig2 * sig2 * sig2;
              const T e4 = nbok ? p2i * pj2 : T(0);
              ljb = e4 * sig6;
              lja = ljb * sig6;
              sigma = (ljb > T(1.0e-6)) ? sig : T(0);
            }
            clash_pair<T, FORCES>(P::sqrt(r2), qnb, lja, ljb, sigma, cd, ca, cb, cratio, e_qq, e_lj,
                                  fmag);
          }
          else {
            const T invr = P::rsqrt(r2);
            const T invr2 = invr * invr;
            const T eqq = qnb * invr;
            e_qq += eqq;
            if (LJTABLE) {
              const double2 ab = ljrow[static_cast<int>(row_ror<K>(p1j))];
              const T lja = nbok ? static_cast<T>(ab.x) : T(0);
              const T ljb = nbok ? static_cast<T>(ab.y) : T(0);
              const T invr4 = invr2 * invr2;
              e_lj += ((lja * invr4 * invr4) - (ljb * invr2)) * invr4;
              if (FORCES) {
                fmag += ((T(6) * ljb) - (T(12) * lja * invr4 * invr2)) * invr4 * invr4;
              }
            }
            else {
              const T sig = row_ror<K>(p1j) + p1i;
              const T s2 = sig * sig * invr2;
              const T s6 = s2 * s2 * s2;
              const T e4c = row_ror<K>(p2j) * p2i;
              const T e4 = nbok ? e4c : T(0);
              const T e4s6 = e4 * s6;
              e_lj += e4s6 * (s6 - T(1));
              if (FORCES) {
                fmag += e4s6 * (T(6) - (T(12) * s6)) * invr2;
              }
            }
            if (FORCES) {
              fmag -= eqq * invr2;
            }
          }
          if (FORCES) {
            const T fdx = fmag * dx;
            const T fdy = fmag * dy;
            const T fdz = fmag * dz;
            fxi += fdx;
            fyi += fdy;
            fzi += fdz;
            fxj = row_rotate(fxj) - fdx;
            fyj = row_rotate(fyj) - fdy;
            fzj = row_rotate(fzj) - fdz;
          }
        };
        unrolled_steps(step, std::make_integer_sequence<int, kGroup>{});
        vbits >>= kGroup;
        nbbits >>= kGroup;
        xj = row_ror<kGroup>(xj);
        yj = row_ror<kGroup>(yj);
        zj = row_ror<kGroup>(zj);
        qj = row_ror<kGroup>(qj);
        p1j = row_ror<kGroup>(p1j);
        p2j = row_ror<kGroup>(p2j);
        if (GB) {
          Rj = row_ror<kGroup>(Rj);
          iRj = row_ror<kGroup>(iRj);
        }
      }
    };
    // The kappa choice is made here too, so that no branch splits the steps.
    const bool clean = wave_all(!tt.diag && tt.ni == kTile && tt.nj == kTile && tk.z < 0);
    if (GB && use_kappa) {
      if (clean) {
        run(std::false_type{}, std::true_type{});
      }
      else {
        run(std::true_type{}, std::true_type{});
      }
    }
    else if (clean) {
      run(std::false_type{}, std::false_type{});
    }
    else {
      run(std::true_type{}, std::false_type{});
    }
    if (FORCES) {
      if (tt.ivalid) {
        lds_add(&L.fx[tt.i], P::fixed(fxi * frc_scale));
        lds_add(&L.fy[tt.i], P::fixed(fyi * frc_scale));
        lds_add(&L.fz[tt.i], P::fixed(fzi * frc_scale));
        if (GB) {
          lds_add(&L.dedr[tt.i], P::fixed(Ri * dri * dedr_scale));
        }
      }
      // After the last step the j sums sit nsteps - 1 lanes right of the lanes that loaded them.
      const int jl_home = tt.jl_home();
      if (jl_home < tt.nj) {
        const int j = tt.jb + jl_home;
        lds_add(&L.fx[j], P::fixed(fxj * frc_scale));
        lds_add(&L.fy[j], P::fixed(fyj * frc_scale));
        lds_add(&L.fz[j], P::fixed(fzj * frc_scale));
        if (GB) {
          lds_add(&L.dedr[j], P::fixed(L.rad[j] * drj * dedr_scale));
        }
      }
    }
    const long long vqq = row_sum(P::fixed(e_qq * nrg_scale));
    const long long vlj = row_sum(P::fixed(e_lj * nrg_scale));
    const long long vgb = GB ? row_sum(P::fixed(e_gb * nrg_scale)) : 0LL;
    if (il == 0) {
      lds_add(&L.nrg[tt.slot * 16 + kElec], vqq);
      lds_add(&L.nrg[tt.slot * 16 + kVdw], vlj);
      if (GB) {
        lds_add(&L.nrg[tt.slot * 16 + kGb], vgb);
      }
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------------------------------------
// OBC chain rule on the accumulated radius derivatives, src/Potential/nonbonded_potential.tpp:
// 985-1017.  The HCT model uses the accumulated values as they stand.
template <typename T>
__device__ void stage_chain(StageCtx<T> &c, const MMKit &k) {
  WuLds<T> &L = c.L;
  for (int a = c.tid; a < c.wu.lds_atoms; a += c.nthr) {
    const int g = L.gidx[a];
    if (g < 0) {
      L.sd[a] = T(0);
      continue;
    }
    double sdi = static_cast<double>(L.dedr[a]) * ldexp(1.0, -kDedrBits);
    if (k.igb != 1) {
      // The radii stage left dR/dpsi in the psi slot.
      sdi *= __longlong_as_double(L.psi[a]);
    }
    L.sd[a] = static_cast<T>(sdi);
  }
  __syncthreads();
}

// ------------------------------------------------------------------------------------------
// Near-field radius-derivative forces of the first `count` queued pairs, one pair per lane.
template <typename T>
__device__ __forceinline__ void deriv_drain(StageCtx<T> &c, const unsigned *q, int count,
                                            int lane, int frc_bits) {
  WuLds<T> &L = c.L;
  if (lane < count) {
    const NearPair<T> p(c, q[lane]);
    const T invr2 = p.invr * p.invr;
    T fmag = T(0);
    if (p.near_i) {
      fmag += dpsi_near(p.r, p.r2, p.invr, invr2, L.rho[p.i], L.irho[p.i], L.scr[p.j]) *
              L.sd[p.i];
    }
    if (p.near_j) {
      fmag += dpsi_near(p.r, p.r2, p.invr, invr2, L.rho[p.j], L.irho[p.j], L.scr[p.i]) *
              L.sd[p.j];
    }
    const long long fx = fixed_contrib(fmag * p.dx, frc_bits);
    const long long fy = fixed_contrib(fmag * p.dy, frc_bits);
    const long long fz = fixed_contrib(fmag * p.dz, frc_bits);
    lds_add(&L.fx[p.i], -fx);
    lds_add(&L.fy[p.i], -fy);
    lds_add(&L.fz[p.i], -fz);
    lds_add(&L.fx[p.j], fx);
    lds_add(&L.fy[p.j], fy);
    lds_add(&L.fz[p.j], fz);
  }
}

// Forces from the Born-radius derivatives, src/Potential/nonbonded_potential.tpp:1018-1292,
// with the same far-field / queued near-field split and step groups as stage_psi.  Far-field
// forces are summed per lane in the working precision and converted to fixed precision once per
// task.
template <typename T>
// Always inlined: kept whole inside the kernels that evaluate forces.
__device__ __forceinline__ void stage_deriv(StageCtx<T> &c, const MMKit &k) {
  using P = Prec<T>;
  WuLds<T> &L = c.L;
  const int il = c.tid & 15;
  const int lane = c.tid & 63;
  const int rlane4 = (63 - lane) << 2;
  // Wavefront-uniform, so the queue's base address stays in scalar registers.
  const int wave = __builtin_amdgcn_readfirstlane(c.tid >> 6);
  unsigned *q = L.nearq + (wave * kNearQueueStride);
  int qn = 0;
  const T frc_scale = T(ldexp(1.0, k.frc_bits));
  const int quads = (c.wu.task_count + 3) >> 2;
  for (int item = next_item(&L.work[1], lane); item < quads; item = next_item(&L.work[1], lane)) {
    const TileTask tt(wave_task(c, k, (item << 2) + ((c.tid >> 4) & 3)), il);
    T xi, yi, zi, xj, yj, zj;
    tile_coords(c, tt, xi, yi, zi, xj, yj, zj);
    const T si = L.scr[tt.i], sdi = L.sd[tt.i];
    const T si2 = si * si;
    const T si3 = si2 * si;
    const T sj = L.scr[tt.j0];
    T sj2 = sj * sj;
    T sj3 = sj2 * sj;
    T sdj = L.sd[tt.j0];
    const unsigned ei = static_cast<unsigned>(tt.i);
    int ej = tt.j0 << 9;
    T fxi = T(0), fyi = T(0), fzi = T(0);
    T gxj = T(0), gyj = T(0), gzj = T(0);
    unsigned vbits = tt.valid;
    // Wavefronts of four off-diagonal tiles run without a pair mask, as in stage_psi.
    auto run = [&](auto mc) {
      constexpr bool MASKED = decltype(mc)::value;
      const int nsteps = MASKED ? tt.nsteps : kTile;
      for (int g = 0; g < nsteps; g += kStepGroup) {
        auto step = [&](auto kc) {
          constexpr int K = decltype(kc)::value;
          if (MASKED && g + K >= nsteps) {
            return;
          }
          // Absent pairs are put far away, as in stage_psi.
          const bool valid = !MASKED || (((vbits >> K) & 1u) != 0u);
          const T dx = row_ror<K>(xj) - xi;
          const T dy = row_ror<K>(yj) - yi;
          const T dz = row_ror<K>(zj) - zi;
          const T r2c = (dx * dx) + (dy * dy) + (dz * dz);
          const T r2 = valid ? r2c : FarAway<T>::kR2;
          const T invr2 = P::rcp(r2);
          const T invr6 = invr2 * invr2 * invr2;
          const T x_i = row_ror<K>(sj2) * invr2;
          const T x_j = si2 * invr2;
          const bool far_i = x_i < T(0.0625);
          const bool far_j = x_j < T(0.0625);
          // Both series are evaluated on every lane and selected, which keeps the step free of
          // branches.
          const T ti = dpsi_series(x_i) * (row_ror<K>(sj3) * (invr6 * sdi));
          const T tj = dpsi_series(x_j) * (row_ror<K>(sdj) * (invr6 * si3));
          const T fmag = (far_i ? ti : T(0)) + (far_j ? tj : T(0));
          const T fdx = fmag * dx;
          const T fdy = fmag * dy;
          const T fdz = fmag * dz;
          fxi -= fdx;
          fyi -= fdy;
          fzi -= fdz;
          gxj = row_rotate(gxj) + fdx;
          gyj = row_rotate(gyj) + fdy;
          gzj = row_rotate(gzj) + fdz;
          nearq_push(q, qn, !far_i, !far_j,
                     ei | static_cast<unsigned>(row_ror<K>(ej)) | (far_i ? 0u : kNearI) |
                         (far_j ? 0u : kNearJ),
                     rlane4);
        };
        unrolled_steps(step, std::integer_sequence<int, 0, 1, 2, 3>{});
        settle(fxi);
        settle(fyi);
        settle(fzi);
        vbits >>= kStepGroup;
        xj = row_ror<kStepGroup>(xj);
        yj = row_ror<kStepGroup>(yj);
        zj = row_ror<kStepGroup>(zj);
        sj2 = row_ror<kStepGroup>(sj2);
        sj3 = row_ror<kStepGroup>(sj3);
        sdj = row_ror<kStepGroup>(sdj);
        ej = row_ror<kStepGroup>(ej);
        while (qn >= 64) {
          deriv_drain(c, q + (qn - 64), 64, lane, k.frc_bits);
          qn -= 64;
        }
      }
    };
    if (wave_all(!tt.diag)) {
      run(std::false_type{});
    }
    else {
      run(std::true_type{});
    }
    if (tt.ivalid) {
      lds_add(&L.fx[tt.i], P::fixed(fxi * frc_scale));
      lds_add(&L.fy[tt.i], P::fixed(fyi * frc_scale));
      lds_add(&L.fz[tt.i], P::fixed(fzi * frc_scale));
    }
    const int jl_home = tt.jl_home();
    if (jl_home < tt.nj) {
      const int j = tt.jb + jl_home;
      lds_add(&L.fx[j], P::fixed(gxj * frc_scale));
      lds_add(&L.fy[j], P::fixed(gyj * frc_scale));
      lds_add(&L.fz[j], P::fixed(gzj * frc_scale));
    }
  }
  deriv_drain(c, q, qn, lane, k.frc_bits);
  __syncthreads();
}

// ------------------------------------------------------------------------------------------
// Valence terms.  They are O(N) work, so both precision models evaluate them in double
// precision: a single-precision r - r0 or theta - theta0 near equilibrium loses most of its
// significant bits.  Displacements come from the LDS coordinates (exact in the double-precision
// model; 2^-24 Angstrom steps in the single-precision model, far finer than the bits such a
// difference keeps in single precision).
// ------------------------------------------------------------------------------------------
struct Vec3 {
  double x, y, z;
};
__device__ __forceinline__ Vec3 cross(const Vec3 &a, const Vec3 &b) {
  return {(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
__device__ __forceinline__ double dot(const Vec3 &a, const Vec3 &b) {
  return (a.x * b.x) + (a.y * b.y) + (a.z * b.z);
}

struct Vec3f {
  float x, y, z;
};
__device__ __forceinline__ Vec3f cross(const Vec3f &a, const Vec3f &b) {
  return {(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
__device__ __forceinline__ float dot(const Vec3f &a, const Vec3f &b) {
  return (a.x * b.x) + (a.y * b.y) + (a.z * b.z);
}

template <typename C> struct ValenceCtxT {
  const C *cx, *cy, *cz;
  const int *gidx;
  double crd_scale;
  float crd_scale_f;
  long long *fx, *fy, *fz;
  // Absolute position (A) of LDS atom a of a system whose coordinate origin is org (fixed
  // precision at inv_gpos).
  __device__ __forceinline__ Vec3 absolute(int a, const long long *org, double inv_gpos) const {
    return {(static_cast<double>(org[0]) * inv_gpos) + (static_cast<double>(cx[a]) * crd_scale),
            (static_cast<double>(org[1]) * inv_gpos) + (static_cast<double>(cy[a]) * crd_scale),
            (static_cast<double>(org[2]) * inv_gpos) + (static_cast<double>(cz[a]) * crd_scale)};
  }
  // Single-precision displacement: the LDS coordinates' difference is exact, then rounded once.
  __device__ __forceinline__ Vec3f dispf(int from, int to) const {
    return {static_cast<float>(cx[to] - cx[from]) * crd_scale_f,
            static_cast<float>(cy[to] - cy[from]) * crd_scale_f,
            static_cast<float>(cz[to] - cz[from]) * crd_scale_f};
  }
  __device__ __forceinline__ Vec3 disp(int from, int to) const {
    return {(static_cast<double>(cx[to]) - static_cast<double>(cx[from])) * crd_scale,
            (static_cast<double>(cy[to]) - static_cast<double>(cy[from])) * crd_scale,
            (static_cast<double>(cz[to]) - static_cast<double>(cz[from])) * crd_scale};
  }
  __device__ __forceinline__ void add(int a, long long vx, long long vy, long long vz) const {
    lds_add(&fx[a], vx);
    lds_add(&fy[a], vy);
    lds_add(&fz[a], vz);
  }
};

__device__ __forceinline__ long long fix64(double v) { return round_to_ll(v); }

// Harmonic bond, src/Potential/valence_potential.tpp:9-96.
template <bool FORCES, typename V>
__device__ __forceinline__ double eval_bond(const V &v, int i, int j, double kb,
                                            double r0, double fscale) {
  const Vec3 d = v.disp(i, j);
  const double dr = sqrt(dot(d, d));
  const double dl = dr - r0;
  if (FORCES) {
    const double fmag = 2.0 * kb * dl / dr;
    const long long ix = fix64(fmag * d.x * fscale);
    const long long iy = fix64(fmag * d.y * fscale);
    const long long iz = fix64(fmag * d.z * fscale);
    v.add(i, ix, iy, iz);
    v.add(j, -ix, -iy, -iz);
  }
  return kb * dl * dl;
}

// Harmonic angle, src/Potential/valence_potential.tpp:170-287.
template <bool FORCES, typename V>
__device__ __forceinline__ double eval_angle(const V &v, int i, int j, int kk,
                                             double ka, double th0, double fscale) {
  const Vec3 ba = v.disp(j, i);
  const Vec3 bc = v.disp(j, kk);
  const double mgba = dot(ba, ba);
  const double mgbc = dot(bc, bc);
  const double invbabc = 1.0 / sqrt(mgba * mgbc);
  double costheta = dot(ba, bc) * invbabc;
  costheta = (costheta < -1.0) ? -1.0 : ((costheta > 1.0) ? 1.0 : costheta);
  const double theta = acos(costheta);
  const double dtheta = theta - th0;
  if (FORCES) {
    const double dA = -2.0 * ka * dtheta / sqrt(1.0 - (costheta * costheta));
    const double sqba = dA / mgba;
    const double sqbc = dA / mgbc;
    const double mbabc = dA * invbabc;
    const long long iax = fix64(((bc.x * mbabc) - (costheta * ba.x * sqba)) * fscale);
    const long long iay = fix64(((bc.y * mbabc) - (costheta * ba.y * sqba)) * fscale);
    const long long iaz = fix64(((bc.z * mbabc) - (costheta * ba.z * sqba)) * fscale);
    const long long icx = fix64(((ba.x * mbabc) - (costheta * bc.x * sqbc)) * fscale);
    const long long icy = fix64(((ba.y * mbabc) - (costheta * bc.y * sqbc)) * fscale);
    const long long icz = fix64(((ba.z * mbabc) - (costheta * bc.z * sqbc)) * fscale);
    v.add(i, -iax, -iay, -iaz);
    v.add(j, iax + icx, iay + icy, iaz + icz);
    v.add(kk, -icx, -icy, -icz);
  }
  return ka * dtheta * dtheta;
}

// Cosine dihedral, src/Potential/valence_potential.tpp:359-578: E = A (1 + cos(n theta - phase))
// with the force factor A n sin(n theta - phase).  The reference recovers the signed angle
// theta with an arccosine (src/Math/rounding.tpp:24-59) and then takes the sine and cosine of
// n theta - phase.  For the integer periodicities used here both functions follow exactly from
// cos(theta) and sin(theta) by the multiple-angle expansion (cos theta + i sin theta)^n, and the
// phase enters through its precomputed cosine and sine.  cos(theta) is the normalized dot
// product of the two plane normals and the signed sin(theta) their cross product projected on
// the central bond (the reference's sign convention for theta), so no inverse or forward
// trigonometric function is evaluated.
template <bool FORCES, typename V>
__device__ __forceinline__ double eval_dihedral(const V &v, int i, int j, int kk, int l,
                                                const DiheParam &dp, double fscale) {
  const Vec3 ab = v.disp(i, j);
  const Vec3 bc = v.disp(j, kk);
  const Vec3 cd = v.disp(kk, l);
  Vec3 crabbc = cross(ab, bc);
  Vec3 crbccd = cross(bc, cd);
  const double inv12 = 1.0 / sqrt(dot(crabbc, crabbc) * dot(crbccd, crbccd));
  const double mgbc = sqrt(dot(bc, bc));
  const Vec3 scr = cross(crabbc, crbccd);
  const double cth = dot(crabbc, crbccd) * inv12;
  const double sth = dot(scr, bc) * inv12 / mgbc;
  double cn = 1.0, sn = 0.0;
  for (int p = 0; p < dp.period; p++) {
    const double t = (cn * cth) - (sn * sth);
    sn = (sn * cth) + (cn * sth);
    cn = t;
  }
  const double cos_sangle = (cn * dp.cos_phase) + (sn * dp.sin_phase);
  if (FORCES) {
    const double sin_sangle = (sn * dp.cos_phase) - (cn * dp.sin_phase);
    const double fr = dp.amp * static_cast<double>(dp.period) * sin_sangle;
    const double mgab = sqrt(dot(ab, ab));
    const double mgcd = sqrt(dot(cd, cd));
    const double invab = 1.0 / mgab;
    const double invbc = 1.0 / mgbc;
    const double invcd = 1.0 / mgcd;
    const double cosb = -dot(ab, bc) * invab * invbc;
    const double cosc = -dot(bc, cd) * invbc * invcd;
    // src/Constants/symbol_values.h:72-82 (asymptotic_to_one_lf, inverse_one_minus_asymptote)
    const double isinb2 = (cosb * cosb < 0.99999904632568359375) ? fr / (1.0 - (cosb * cosb)) :
                                                                   fr * 1048576.0;
    const double isinc2 = (cosc * cosc < 0.99999904632568359375) ? fr / (1.0 - (cosc * cosc)) :
                                                                   fr * 1048576.0;
    const double invabc = invab * invbc;
    const double invbcd = invbc * invcd;
    crabbc = {crabbc.x * invabc, crabbc.y * invabc, crabbc.z * invabc};
    crbccd = {crbccd.x * invbcd, crbccd.y * invbcd, crbccd.z * invbcd};
    const double fa = -invab * isinb2;
    const double fb1 = (mgbc - (mgab * cosb)) * invabc * isinb2;
    const double fb2 = cosc * invbc * isinc2;
    const double fd = -invcd * isinc2;
    const long long fix = fix64(crabbc.x * fa * fscale);
    const long long fiy = fix64(crabbc.y * fa * fscale);
    const long long fiz = fix64(crabbc.z * fa * fscale);
    v.add(i, fix, fiy, fiz);
    const long long flx = fix64(-fd * crbccd.x * fscale);
    const long long fly = fix64(-fd * crbccd.y * fscale);
    const long long flz = fix64(-fd * crbccd.z * fscale);
    v.add(l, flx, fly, flz);
    const long long fjx = fix64(((fb1 * crabbc.x) - (fb2 * crbccd.x)) * fscale);
    const long long fjy = fix64(((fb1 * crabbc.y) - (fb2 * crbccd.y)) * fscale);
    const long long fjz = fix64(((fb1 * crabbc.z) - (fb2 * crbccd.z)) * fscale);
    v.add(j, fjx, fjy, fjz);
    v.add(kk, -(fix + fjx + flx), -(fiy + fjy + fly), -(fiz + fjz + flz));
  }
  return dp.amp * (1.0 + cos_sangle);
}

// Attenuated 1:4 pair, src/Potential/valence_potential.tpp:1288-1419; with clash damping on,
// the damped potentials of src/Potential/valence_potential.cui:930-1040, whose sigma comes from
// the attenuated coefficients.
template <bool FORCES, bool CLASH, typename V>
__device__ __forceinline__ void eval_14(const V &v, const MMKit &k, bool lj_table, int2 lj,
                                        int i, int l, double inv_scee, double inv_scnb,
                                        double fscale, bool clash_on, double cd, double cratio,
                                        double &ele, double &vdw) {
  const int gi = v.gidx[i], gl = v.gidx[l];
  const Vec3 d = v.disp(i, l);
  const double r2 = dot(d, d);
  const double qiqj = k.q[gi] * k.q[gl] * inv_scee;
  double lja, ljb, sigma_lb = 0.0;
  if (lj_table) {
    const double2 ab = k.ljab14[lj.x + (k.lj_type[gi] * lj.y) + k.lj_type[gl]];
    lja = ab.x * inv_scnb;
    ljb = ab.y * inv_scnb;
  }
  else {
    sigma_lb = 0.5 * (k.ljp1[gi] + k.ljp1[gl]);
    const double sig2 = sigma_lb * sigma_lb;
    const double sig6 = sig2 * sig2 * sig2;
    ljb = k.ljp2[gi] * k.ljp2[gl] * inv_scnb * sig6;
    lja = ljb * sig6;
  }
  double fmag = 0.0;
  if (CLASH && clash_on) {
    ele = 0.0;
    vdw = 0.0;
    const double sigma = (ljb > 1.0e-6) ? (lj_table ? sqrt(cbrt(lja / ljb)) : sigma_lb) : 0.0;
    const double ca = -0.5 / (cd * cd * (cd + 1.0));
    const double cb = (1.0 / cd) - (ca * (cd + 1.0) * (cd + 1.0));
    clash_pair<double, FORCES>(sqrt(r2), qiqj, lja, ljb, sigma, cd, ca, cb, cratio, ele, vdw,
                               fmag);
  }
  else {
    const double invr2 = 1.0 / r2;
    const double invr = sqrt(invr2);
    const double invr4 = invr2 * invr2;
    ele = qiqj * invr;
    vdw = (lja * invr4 * invr4 * invr4) - (ljb * invr4 * invr2);
    if (FORCES) {
      fmag = ((6.0 * ljb) - (12.0 * lja * invr4 * invr2)) * invr4 * invr4;
      fmag -= qiqj * invr * invr2;
    }
  }
  if (FORCES) {
    const long long ix = fix64(fmag * d.x * fscale);
    const long long iy = fix64(fmag * d.y * fscale);
    const long long iz = fix64(fmag * d.z * fscale);
    v.add(i, ix, iy, iz);
    v.add(l, -ix, -iy, -iz);
  }
}

// Single-precision forms of the cosine dihedral and the attenuated 1:4 pair for the
// single-precision model, as the reference evaluates them there (valence_potential.cui:540-640,
// 930-1040 with TCALC = float).  Their energies carry no cancellation (a cosine series and
// pair potentials at 1:4 distances), unlike r - r0 and theta - theta0 of the bonds and angles,
// which stay in double precision.  Guards: asymptotic_to_one_f and
// inverse_one_minus_asymptote_f (src/Constants/symbol_values.h:72-82).
template <bool FORCES, typename V>
__device__ __forceinline__ float eval_dihedral_f(const V &v, int i, int j, int kk, int l,
                                                 const DiheParam &dp, int frc_bits) {
  using P = Prec<float>;
  const Vec3f ab = v.dispf(i, j);
  const Vec3f bc = v.dispf(j, kk);
  const Vec3f cd = v.dispf(kk, l);
  Vec3f crabbc = cross(ab, bc);
  Vec3f crbccd = cross(bc, cd);
  const float inv12 = P::rsqrt(dot(crabbc, crabbc) * dot(crbccd, crbccd));
  const float invbc = P::rsqrt(dot(bc, bc));
  const Vec3f scr = cross(crabbc, crbccd);
  const float cth = dot(crabbc, crbccd) * inv12;
  const float sth = dot(scr, bc) * inv12 * invbc;
  float cn = 1.0f, sn = 0.0f;
  for (int p = 0; p < dp.period; p++) {
    const float t = (cn * cth) - (sn * sth);
    sn = (sn * cth) + (cn * sth);
    cn = t;
  }
  const float cph = static_cast<float>(dp.cos_phase);
  const float sph = static_cast<float>(dp.sin_phase);
  const float amp = static_cast<float>(dp.amp);
  const float cos_sangle = (cn * cph) + (sn * sph);
  if (FORCES) {
    const float sin_sangle = (sn * cph) - (cn * sph);
    const float fr = amp * static_cast<float>(dp.period) * sin_sangle;
    const float invab = P::rsqrt(dot(ab, ab));
    const float invcd = P::rsqrt(dot(cd, cd));
    const float mgab = dot(ab, ab) * invab;
    const float mgbc = dot(bc, bc) * invbc;
    const float cosb = -dot(ab, bc) * invab * invbc;
    const float cosc = -dot(bc, cd) * invbc * invcd;
    constexpr float kAsym = 0.99999904632568359375f;
    constexpr float kInvAsym = 1048576.0f;
    const float isinb2 = (cosb * cosb < kAsym) ? fr * P::rcp(1.0f - (cosb * cosb)) : fr * kInvAsym;
    const float isinc2 = (cosc * cosc < kAsym) ? fr * P::rcp(1.0f - (cosc * cosc)) : fr * kInvAsym;
    const float invabc = invab * invbc;
    const float invbcd = invbc * invcd;
    crabbc = {crabbc.x * invabc, crabbc.y * invabc, crabbc.z * invabc};
    crbccd = {crbccd.x * invbcd, crbccd.y * invbcd, crbccd.z * invbcd};
    const float fa = -invab * isinb2;
    const float fb1 = (mgbc - (mgab * cosb)) * invabc * isinb2;
    const float fb2 = cosc * invbc * isinc2;
    const float fd = -invcd * isinc2;
    const long long fix = fixed_contrib(crabbc.x * fa, frc_bits);
    const long long fiy = fixed_contrib(crabbc.y * fa, frc_bits);
    const long long fiz = fixed_contrib(crabbc.z * fa, frc_bits);
    v.add(i, fix, fiy, fiz);
    const long long flx = fixed_contrib(-fd * crbccd.x, frc_bits);
    const long long fly = fixed_contrib(-fd * crbccd.y, frc_bits);
    const long long flz = fixed_contrib(-fd * crbccd.z, frc_bits);
    v.add(l, flx, fly, flz);
    const long long fjx = fixed_contrib((fb1 * crabbc.x) - (fb2 * crbccd.x), frc_bits);
    const long long fjy = fixed_contrib((fb1 * crabbc.y) - (fb2 * crbccd.y), frc_bits);
    const long long fjz = fixed_contrib((fb1 * crabbc.z) - (fb2 * crbccd.z), frc_bits);
    v.add(j, fjx, fjy, fjz);
    v.add(kk, -(fix + fjx + flx), -(fiy + fjy + fly), -(fiz + fjz + flz));
  }
  return amp * (1.0f + cos_sangle);
}

template <bool FORCES, bool CLASH, typename V>
__device__ __forceinline__ void eval_14_f(const V &v, const MMKit &k, bool lj_table, int2 lj,
                                          int i, int l, float inv_scee, float inv_scnb,
                                          int frc_bits, bool clash_on, float cd, float cratio,
                                          float &ele, float &vdw) {
  using P = Prec<float>;
  const int gi = v.gidx[i], gl = v.gidx[l];
  const Vec3f d = v.dispf(i, l);
  const float r2 = dot(d, d);
  const float qiqj = static_cast<float>(k.q[gi] * k.q[gl]) * inv_scee;
  float lja, ljb, sigma_lb = 0.0f;
  if (lj_table) {
    const double2 ab = k.ljab14[lj.x + (k.lj_type[gi] * lj.y) + k.lj_type[gl]];
    lja = static_cast<float>(ab.x) * inv_scnb;
    ljb = static_cast<float>(ab.y) * inv_scnb;
  }
  else {
    sigma_lb = 0.5f * static_cast<float>(k.ljp1[gi] + k.ljp1[gl]);
    const float sig2 = sigma_lb * sigma_lb;
    const float sig6 = sig2 * sig2 * sig2;
    ljb = static_cast<float>(k.ljp2[gi] * k.ljp2[gl]) * inv_scnb * sig6;
    lja = ljb * sig6;
  }
  float fmag = 0.0f;
  if (CLASH && clash_on) {
    ele = 0.0f;
    vdw = 0.0f;
    const float sigma = (ljb > 1.0e-6f) ? (lj_table ? P::sqrt(P::cbrt(lja / ljb)) : sigma_lb)
                                        : 0.0f;
    const float ca = -0.5f / (cd * cd * (cd + 1.0f));
    const float cb = (1.0f / cd) - (ca * (cd + 1.0f) * (cd + 1.0f));
    clash_pair<float, FORCES>(P::sqrt(r2), qiqj, lja, ljb, sigma, cd, ca, cb, cratio, ele, vdw,
                              fmag);
  }
  else {
    const float invr = P::rsqrt(r2);
    const float invr2 = invr * invr;
    const float invr4 = invr2 * invr2;
    ele = qiqj * invr;
    vdw = (lja * invr4 * invr4 * invr4) - (ljb * invr4 * invr2);
    if (FORCES) {
      fmag = ((6.0f * ljb) - (12.0f * lja * invr4 * invr2)) * invr4 * invr4;
      fmag -= qiqj * invr * invr2;
    }
  }
  if (FORCES) {
    const long long ix = fixed_contrib(fmag * d.x, frc_bits);
    const long long iy = fixed_contrib(fmag * d.y, frc_bits);
    const long long iz = fixed_contrib(fmag * d.z, frc_bits);
    v.add(i, ix, iy, iz);
    v.add(l, -ix, -iy, -iz);
  }
}

// ------------------------------------------------------------------------------------------
// Flat-bottom restraints, src/Restraints/restraint_util.tpp:8-77 and
// src/Potential/valence_potential.tpp:1555-2088 (non-periodic systems: no imaging), in double
// precision like the other valence terms.
// ------------------------------------------------------------------------------------------
// Weights of a restraint's initial and final parameters at `step`.
__device__ __forceinline__ double2 restraint_mixture(int step, int init_step, int final_step) {
  if (step < init_step) {
    return make_double2(0.0, 0.0);
  }
  if (init_step == final_step) {
    return make_double2(1.0, 0.0);
  }
  if (step < final_step) {
    const double wslide = static_cast<double>(step - init_step) /
                          static_cast<double>(final_step - init_step);
    return make_double2(1.0 - wslide, wslide);
  }
  return make_double2(0.0, 1.0);
}

// Stiffness, displacement beyond the flat bottom and penalty at coordinate value dr.
__device__ __forceinline__ double3 restraint_delta(const stormm_restraint &t, double2 w,
                                                   double dr) {
  const double r1 = (w.x * t.init_r[0]) + (w.y * t.final_r[0]);
  const double r2 = (w.x * t.init_r[1]) + (w.y * t.final_r[1]);
  const double r3 = (w.x * t.init_r[2]) + (w.y * t.final_r[2]);
  const double r4 = (w.x * t.init_r[3]) + (w.y * t.final_r[3]);
  const double k2 = (w.x * t.init_k[0]) + (w.y * t.final_k[0]);
  const double k3 = (w.x * t.init_k[1]) + (w.y * t.final_k[1]);
  if (dr < r1) {
    const double dl = r1 - r2;
    return make_double3(k2, dl, k2 * ((dl * dl) + (2.0 * dl * (dr - r1))));
  }
  if (dr < r2) {
    const double dl = dr - r2;
    return make_double3(k2, dl, k2 * dl * dl);
  }
  if (dr < r3) {
    return make_double3(0.0, 0.0, 0.0);
  }
  if (dr < r4) {
    const double dl = dr - r3;
    return make_double3(k3, dl, k3 * dl * dl);
  }
  const double dl = r4 - r3;
  return make_double3(k3, dl, k3 * ((dl * dl) + (2.0 * dl * (dr - r4))));
}

// Minimum image of an angle over 2 pi, src/Structure/local_arrangement.tpp:42-58.
__device__ __forceinline__ double image_angle(double x) {
  constexpr double kTwoPi = 6.283185307179586476925286766559;
  double xf = x / kTwoPi;
  xf -= ((xf >= 0.5) ? ceil(xf - 0.5) : 0.0) + ((xf < -0.5) ? floor(xf + 0.5) : 0.0);
  return (xf - ((xf >= 0.5) ? 1.0 : 0.0)) * kTwoPi;
}

// One restraint of the system at LDS base `base`; returns its penalty.
template <bool FORCES, typename V>
__device__ double eval_restraint(const V &v, const stormm_restraint &t, int base,
                                 const long long *org, double inv_gpos, int step,
                                 double fscale) {
  const double2 w = restraint_mixture(step, t.init_step, t.final_step);
  const int i = base + t.atoms[0], j = base + t.atoms[1];
  const int kk = base + t.atoms[2], l = base + t.atoms[3];
  if (t.kind == STORMM_RESTRAINT_POSITION) {
    const Vec3 p = v.absolute(i, org, inv_gpos);
    const Vec3 d = {p.x - ((w.x * t.init_target[0]) + (w.y * t.final_target[0])),
                    p.y - ((w.x * t.init_target[1]) + (w.y * t.final_target[1])),
                    p.z - ((w.x * t.init_target[2]) + (w.y * t.final_target[2]))};
    const double dr = sqrt(dot(d, d));
    const double3 e = restraint_delta(t, w, dr);
    if (FORCES) {
      if (dr < 1.0e-10) {
        // On the target: split the force among the three axes (valence_potential.tpp:1610-1626).
        const long long f = fix64(2.0 * e.x * e.y / sqrt(3.0) * fscale);
        v.add(i, -f, -f, -f);
      }
      else {
        const double fmag = 2.0 * e.x * e.y / dr;
        v.add(i, -fix64(fmag * d.x * fscale), -fix64(fmag * d.y * fscale),
              -fix64(fmag * d.z * fscale));
      }
    }
    return e.z;
  }
  if (t.kind == STORMM_RESTRAINT_DISTANCE) {
    const Vec3 d = v.disp(i, j);
    const double dr = sqrt(dot(d, d));
    const double3 e = restraint_delta(t, w, dr);
    if (FORCES) {
      const double fmag = 2.0 * e.x * e.y / dr;
      const long long fx = fix64(fmag * d.x * fscale);
      const long long fy = fix64(fmag * d.y * fscale);
      const long long fz = fix64(fmag * d.z * fscale);
      v.add(i, fx, fy, fz);
      v.add(j, -fx, -fy, -fz);
    }
    return e.z;
  }
  if (t.kind == STORMM_RESTRAINT_ANGLE) {
    const Vec3 ba = v.disp(j, i);
    const Vec3 bc = v.disp(j, kk);
    const double mgba = dot(ba, ba);
    const double mgbc = dot(bc, bc);
    const double invbabc = 1.0 / sqrt(mgba * mgbc);
    double costheta = dot(ba, bc) * invbabc;
    costheta = (costheta < -1.0) ? -1.0 : ((costheta > 1.0) ? 1.0 : costheta);
    const double3 e = restraint_delta(t, w, acos(costheta));
    if (FORCES) {
      const double dA = -2.0 * e.x * e.y / sqrt(1.0 - (costheta * costheta));
      const double sqba = dA / mgba;
      const double sqbc = dA / mgbc;
      const double mbabc = dA * invbabc;
      const long long iax = fix64(((bc.x * mbabc) - (costheta * ba.x * sqba)) * fscale);
      const long long iay = fix64(((bc.y * mbabc) - (costheta * ba.y * sqba)) * fscale);
      const long long iaz = fix64(((bc.z * mbabc) - (costheta * ba.z * sqba)) * fscale);
      const long long icx = fix64(((ba.x * mbabc) - (costheta * bc.x * sqbc)) * fscale);
      const long long icy = fix64(((ba.y * mbabc) - (costheta * bc.y * sqbc)) * fscale);
      const long long icz = fix64(((ba.z * mbabc) - (costheta * bc.z * sqbc)) * fscale);
      v.add(i, -iax, -iay, -iaz);
      v.add(j, iax + icx, iay + icy, iaz + icz);
      v.add(kk, -icx, -icy, -icz);
    }
    return e.z;
  }
  // Dihedral: the signed angle from the plane normals, imaged to the restraint's centre.
  const Vec3 ab = v.disp(i, j);
  const Vec3 bc = v.disp(j, kk);
  const Vec3 cd = v.disp(kk, l);
  Vec3 crabbc = cross(ab, bc);
  Vec3 crbccd = cross(bc, cd);
  double costheta = dot(crabbc, crbccd) / sqrt(dot(crabbc, crabbc) * dot(crbccd, crbccd));
  const Vec3 scr = cross(crabbc, crbccd);
  costheta = (costheta < -1.0) ? -1.0 : ((costheta > 1.0) ? 1.0 : costheta);
  double theta = (dot(scr, bc) > 0.0) ? acos(costheta) : -acos(costheta);
  const double midpoint = 0.5 * ((w.x * (t.init_r[1] + t.init_r[2])) +
                                 (w.y * (t.final_r[1] + t.final_r[2])));
  theta += image_angle(theta - midpoint) - (theta - midpoint);
  const double3 e = restraint_delta(t, w, theta);
  if (FORCES) {
    const double fr = -2.0 * e.x * e.y;
    const double mgab = sqrt(dot(ab, ab));
    const double mgbc = sqrt(dot(bc, bc));
    const double mgcd = sqrt(dot(cd, cd));
    const double invab = 1.0 / mgab;
    const double invbc = 1.0 / mgbc;
    const double invcd = 1.0 / mgcd;
    const double cosb = -dot(ab, bc) * invab * invbc;
    const double cosc = -dot(bc, cd) * invbc * invcd;
    // src/Constants/symbol_values.h:72-82 (asymptotic_to_one_lf, inverse_one_minus_asymptote)
    const double isinb2 = (cosb * cosb < 0.99999904632568359375) ? fr / (1.0 - (cosb * cosb)) :
                                                                   fr * 1048576.0;
    const double isinc2 = (cosc * cosc < 0.99999904632568359375) ? fr / (1.0 - (cosc * cosc)) :
                                                                   fr * 1048576.0;
    const double invabc = invab * invbc;
    const double invbcd = invbc * invcd;
    crabbc = {crabbc.x * invabc, crabbc.y * invabc, crabbc.z * invabc};
    crbccd = {crbccd.x * invbcd, crbccd.y * invbcd, crbccd.z * invbcd};
    const double fa = -invab * isinb2;
    const double fb1 = (mgbc - (mgab * cosb)) * invabc * isinb2;
    const double fb2 = cosc * invbc * isinc2;
    const double fd = -invcd * isinc2;
    const long long fix = fix64(crabbc.x * fa * fscale);
    const long long fiy = fix64(crabbc.y * fa * fscale);
    const long long fiz = fix64(crabbc.z * fa * fscale);
    const long long flx = fix64(-fd * crbccd.x * fscale);
    const long long fly = fix64(-fd * crbccd.y * fscale);
    const long long flz = fix64(-fd * crbccd.z * fscale);
    const long long fjx = fix64(((fb1 * crabbc.x) - (fb2 * crbccd.x)) * fscale);
    const long long fjy = fix64(((fb1 * crabbc.y) - (fb2 * crbccd.y)) * fscale);
    const long long fjz = fix64(((fb1 * crabbc.z) - (fb2 * crbccd.z)) * fscale);
    v.add(i, fix, fiy, fiz);
    v.add(l, flx, fly, flz);
    v.add(j, fjx, fjy, fjz);
    v.add(kk, -(fix + fjx + flx), -(fiy + fjy + fly), -(fiz + fjz + flz));
  }
  return e.z;
}

// Valence term t of the work unit (nothing when t is past the unit's last term).
template <typename T, bool FORCES, bool CLASH>
__device__ void valence_term(StageCtx<T> &c, const MMKit &k, int kind, int t) {
  const int *pre = c.L.vpre + (kind * (kMaxWuSystems + 1));
  if (t >= pre[c.wu.sys_count]) {
    return;
  }
  const double fscale = ldexp(1.0, k.frc_bits);
  const double nrg_scale = ldexp(1.0, k.nrg_bits);
  ValenceCtxT<typename Prec<T>::crd_t> v;
  v.cx = c.L.cx;
  v.cy = c.L.cy;
  v.cz = c.L.cz;
  v.gidx = c.L.gidx;
  v.crd_scale = static_cast<double>(c.crd_scale);
  v.crd_scale_f = static_cast<float>(c.crd_scale);
  v.fx = c.L.fx;
  v.fy = c.L.fy;
  v.fz = c.L.fz;
  // The system's terms of this kind, in its template's bonds / angles / dihedrals order.
  const int slot = find_slot(pre, c.wu.sys_count, t);
  const WuSys &ws = c.L.sys[slot];
  const int kind_first = (kind == kVBond) ? 0 :
                         (kind == kVAngle) ? ws.bond_count : ws.bond_count + ws.angl_count;
  const int4 vt = k.vterms[ws.term_first + kind_first + (t - pre[slot])];
  const int i = ws.lds_base + (vt.y & 0xffff), j = ws.lds_base + (vt.y >> 16);
  const int kk = ws.lds_base + (vt.z & 0xffff), l = ws.lds_base + (vt.z >> 16);
  long long *nrg = &c.L.nrg[slot * 16];
  if (kind == kVBond) {
    const double2 p = k.bond_par[vt.w];
    lds_add(&nrg[kBond], fix64(eval_bond<FORCES>(v, i, j, p.x, p.y, fscale) * nrg_scale));
  }
  else if (kind == kVAngle) {
    const double2 p = k.angl_par[vt.w];
    lds_add(&nrg[kAngle], fix64(eval_angle<FORCES>(v, i, j, kk, p.x, p.y, fscale) * nrg_scale));
  }
  else if (std::is_same<T, float>::value) {
    const DiheParam p = k.dihe_par[vt.w];
    const float e = eval_dihedral_f<FORCES>(v, i, j, kk, l, p, k.frc_bits);
    lds_add(&nrg[p.improper ? kImproperDihedral : kProperDihedral],
            fix64(static_cast<double>(e) * nrg_scale));
    const double2 a14 = k.dihe14[vt.w];
    if (a14.x > 0.0) {
      float ele, vdw;
      const bool con = CLASH && (c.L.clash_on[slot] != 0);
      eval_14_f<FORCES, CLASH>(v, k, c.wu.lj_table != 0, make_int2(ws.lj_off, ws.lj_nt), i, l,
                               static_cast<float>(a14.x), static_cast<float>(a14.y), k.frc_bits,
                               con, con ? static_cast<float>(c.L.clash_d[slot]) : 0.0f,
                               con ? static_cast<float>(c.L.clash_r[slot]) : 0.0f, ele, vdw);
      lds_add(&nrg[kElec14], fix64(static_cast<double>(ele) * nrg_scale));
      lds_add(&nrg[kVdw14], fix64(static_cast<double>(vdw) * nrg_scale));
    }
  }
  else {
    const DiheParam p = k.dihe_par[vt.w];
    const double e = eval_dihedral<FORCES>(v, i, j, kk, l, p, fscale);
    lds_add(&nrg[p.improper ? kImproperDihedral : kProperDihedral], fix64(e * nrg_scale));
    const double2 a14 = k.dihe14[vt.w];
    if (a14.x > 0.0) {
      double ele, vdw;
      const bool con = CLASH && (c.L.clash_on[slot] != 0);
      eval_14<FORCES, CLASH>(v, k, c.wu.lj_table != 0, make_int2(ws.lj_off, ws.lj_nt), i, l,
                             a14.x, a14.y, fscale, con,
                             con ? static_cast<double>(c.L.clash_d[slot]) : 0.0,
                             con ? static_cast<double>(c.L.clash_r[slot]) : 0.0, ele, vdw);
      lds_add(&nrg[kElec14], fix64(ele * nrg_scale));
      lds_add(&nrg[kVdw14], fix64(vdw * nrg_scale));
    }
  }
}

// Valence terms on their own (evaluations without pair tiles), in chunks of 64 per wavefront.
template <typename T, bool FORCES, bool CLASH>
__device__ void stage_valence(StageCtx<T> &c, const MMKit &k) {
  const int lane = c.tid & 63;
  const int chunks = valence_chunks(c);
  for (int item = next_item(&c.L.work[2], lane); item < chunks;
       item = next_item(&c.L.work[2], lane)) {
    valence_chunk<T, FORCES, CLASH>(c, k, item, lane);
  }
  __syncthreads();
}

// Restraints of the unit's systems, one thread each, in a stage of their own so that their
// double-precision code stays out of the register budget of the tile loops.
template <typename T, bool FORCES>
__device__ void stage_restraints(StageCtx<T> &c, const MMKit &k) {
  const double fscale = ldexp(1.0, k.frc_bits);
  const double nrg_scale = ldexp(1.0, k.nrg_bits);
  const double inv_gpos = ldexp(1.0, -(k.gpos_bits - c.org_shift));
  ValenceCtxT<typename Prec<T>::crd_t> v;
  v.cx = c.L.cx;
  v.cy = c.L.cy;
  v.cz = c.L.cz;
  v.gidx = c.L.gidx;
  v.crd_scale = static_cast<double>(c.crd_scale);
  v.crd_scale_f = static_cast<float>(c.crd_scale);
  v.fx = c.L.fx;
  v.fy = c.L.fy;
  v.fz = c.L.fz;
  for (int r = c.tid; r < c.wu.rst_count; r += c.nthr) {
    int slot = 0, u = r;
    while (u >= c.L.sys[slot].rst_count) {
      u -= c.L.sys[slot].rst_count;
      slot++;
    }
    const WuSys &ws = c.L.sys[slot];
    const int step = (k.rst_step_sys != nullptr) ?
                         minimizer_step(k.rst_step_sys[ws.sys_index] + k.rst_cycle_offset,
                                        k.rst_move, k.rst_clash_cycles) :
                         c.rst_step;
    const double e = eval_restraint<FORCES>(v, k.rst[ws.rst_first + u], ws.lds_base,
                                            &c.L.org[3 * slot], inv_gpos, step, fscale);
    lds_add(&c.L.nrg[slot * 16 + kRestraint], fix64(e * nrg_scale));
  }
  __syncthreads();
}

// ------------------------------------------------------------------------------------------
// One evaluation of the selected terms on the coordinates currently in global memory, leaving
// forces (FORCES) and energies in the LDS accumulators.  RESTRAINTS = false leaves the restraint
// stage out of the instantiation (for callers that never see restrained work units).
template <typename T, bool FORCES, bool CLASH, bool RESTRAINTS = true>
__device__ __forceinline__ void evaluate_wu(StageCtx<T> &c, const MMKit &k, int terms) {
  const bool lj_table = (c.wu.lj_table != 0);
  stage_load_coords<T, CLASH>(c, k);
  // With GB the valence terms ride along with the Born-integral tiles.
  const bool val = (terms & kTermValence) != 0;
  if (c.gb) {
    stage_psi<T, FORCES, CLASH>(c, k, val ? valence_chunks(c) : 0);
    stage_radii<T, FORCES>(c, k, 0.0);
  }
  if (terms & kTermNonbonded) {
    if (lj_table) {
      if (c.gb) {
        stage_pairs<T, FORCES, true, true, CLASH>(c, k);
      }
      else {
        stage_pairs<T, FORCES, true, false, CLASH>(c, k);
      }
    }
    else {
      if (c.gb) {
        stage_pairs<T, FORCES, false, true, CLASH>(c, k);
      }
      else {
        stage_pairs<T, FORCES, false, false, CLASH>(c, k);
      }
    }
  }
  if (FORCES && c.gb) {
    stage_chain<T>(c, k);
    stage_deriv<T>(c, k);
  }
  if (val && !c.gb) {
    stage_valence<T, FORCES, CLASH>(c, k);
  }
  if (RESTRAINTS && val && c.wu.rst_count > 0) {
    stage_restraints<T, FORCES>(c, k);
  }
}

// ------------------------------------------------------------------------------------------
// Results out: forces added to the global fixed-precision arrays (one owning block per atom)
// and energies added to the ScoreCard rows of the block's systems.
template <typename T, bool FORCES>
__device__ __forceinline__ void stage_store(StageCtx<T> &c, const MMKit &k, bool eval_energy) {
  WuLds<T> &L = c.L;
  if (FORCES && k.xfrc_ovrf != nullptr) {
    // 95-bit forces: the LDS sums scaled to the stored precision, added to (or replacing) the
    // stored values and split into word and overflow word.
    long long *frc[3] = {k.xfrc, k.yfrc, k.zfrc};
    int *ovr[3] = {k.xfrc_ovrf, k.yfrc_ovrf, k.zfrc_ovrf};
    const long long *lds[3] = {L.fx, L.fy, L.fz};
    for (int a = c.tid; a < c.wu.lds_atoms; a += c.nthr) {
      const int g = L.gidx[a];
      if (g >= 0) {
        for (int d = 0; d < 3; d++) {
          __int128 v = static_cast<__int128>(lds[d][a]) << k.frc_store_shift;
          if (!k.store_overwrite) {
            v += i95_value(frc[d][g], ovr[d][g]);
          }
          i95_split(v, &frc[d][g], &ovr[d][g]);
        }
      }
    }
  }
  else if (FORCES) {
    for (int a = c.tid; a < c.wu.lds_atoms; a += c.nthr) {
      const int g = L.gidx[a];
      if (g >= 0) {
        if (k.store_overwrite) {
          k.xfrc[g] = L.fx[a];
          k.yfrc[g] = L.fy[a];
          k.zfrc[g] = L.fz[a];
        }
        else {
          k.xfrc[g] += L.fx[a];
          k.yfrc[g] += L.fy[a];
          k.zfrc[g] += L.fz[a];
        }
      }
    }
  }
  if (eval_energy) {
    for (int e = c.tid; e < c.wu.sys_count * kNrgSlots; e += c.nthr) {
      const int s = e / kNrgSlots;
      const int comp = e - (s * kNrgSlots);
      const long long v = L.nrg[s * 16 + comp];
      long long *dst = &k.scorecard[(static_cast<size_t>(L.sys[s].sys_index) * kScoreStride) +
                                    comp];
      if (k.store_overwrite) {
        *dst = v;
      }
      else if (v != 0) {
        *dst += v;
      }
    }
  }
}

}  // namespace stormm_amd
