// Host side of the C-ABI: AtomGraphSynthesis construction (parameter layout, Lennard-Jones mode
// detection, work-unit packing, tile tasks, exclusion masks and valence term streams), the
// PhaseSpaceSynthesis / ScoreCard / controls objects, and the launch entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <climits>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/stormm_amd.h"
#include "host_common.h"
#include "int95.h"
#include "internal.h"
#include "kernels.h"
#include "plan.h"

using namespace stormm_amd;

// ------------------------------------------------------------------------------------------
// Error reporting
// ------------------------------------------------------------------------------------------
static thread_local std::string g_last_error;

namespace stormm_amd {
int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
}  // namespace stormm_amd

extern "C" const char *stormm_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------------------------------
// Objects
// ------------------------------------------------------------------------------------------
struct stormm_synthesis {
  SynthesisPlan plan;
  DevMem atom_starts, atom_counts, q, ljp1, ljp2, lj_type, ljab, ljab14, rho, scr,
      pbr, alpha, beta, gamma, wu, wu_sys, tasks, masks, vterms, bond_par, angl_par, dihe_par,
      dihe14, status, cycles_run, rms_force, stop_flags, rst;
  MMKit kit;
};

struct stormm_phasespace {
  const stormm_synthesis *ag = nullptr;
  int gpos_bits = 32, frc_bits = 23;
  DevMem xcrd, ycrd, zcrd, xfrc, yfrc, zfrc, xprv, yprv, zprv, xcg, ycg, zcg;
  // Overflow words of the same arrays (value = ovrf 2^63 + word above the plain limits).
  DevMem xcrd_o, ycrd_o, zcrd_o, xfrc_o, yfrc_o, zfrc_o, xprv_o, yprv_o, zprv_o, xcg_o, ycg_o,
      zcg_o;
};

struct stormm_scorecard {
  int nsys = 0, nrg_bits = 26, samples = 0;
  // Samples the time series holds (the reference's default of 16 until it is first used).
  int capacity = 16;
  DevMem inst, running, squared;
  DevMem series;               // [capacity][nsys][32] int64, allocated on first use
  std::vector<int> time_steps; // one per sample
};

struct stormm_mm_controls {
  stormm_minimize_controls mc;
  int step = 0;
};

struct stormm_line_min {
  int nsys = 0;
  DevMem vals;  // 9 * nsys doubles
};

extern "C" int stormm_device_info(int32_t *cu_count, int32_t *lds_bytes_per_cu, char *name,
                                  int32_t name_len) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (cu_count) {
    *cu_count = prop.multiProcessorCount;
  }
  if (lds_bytes_per_cu) {
    *lds_bytes_per_cu = static_cast<int32_t>(prop.maxSharedMemoryPerMultiProcessor);
  }
  if (name && name_len > 0) {
    std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  return STORMM_OK;
}

extern "C" int stormm_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return STORMM_OK;
}

extern "C" int stormm_synthesis_plan(const stormm_topology_batch *batch,
                                     const stormm_launch_shape *shape,
                                     stormm_synthesis_info *info) {
  SynthesisPlan P;
  const int rc = build_plan(batch, shape, P);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (info) {
    info->system_count = P.nsys;
    info->padded_atom_count = P.padded_atoms;
    info->work_unit_count = static_cast<int>(P.wu.size());
    info->task_count = static_cast<int>(P.task_count);
    info->max_wu_atoms = P.maxa;
    info->lj_table_systems = P.lj_table_systems;
    info->pair_count = P.pairs;
    info->valence_term_count = P.vterm_count;
  }
  return STORMM_OK;
}

extern "C" int stormm_synthesis_create(const stormm_topology_batch *batch,
                                       const stormm_launch_shape *shape,
                                       stormm_synthesis **out) {
  if (out == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null output handle");
  }
  std::unique_ptr<stormm_synthesis> ag(new stormm_synthesis());
  int rc = build_plan(batch, shape, ag->plan);
  if (rc != STORMM_OK) {
    return rc;
  }
  SynthesisPlan &P = ag->plan;
  HIP_TRY(ag->atom_starts.upload(P.atom_starts));
  HIP_TRY(ag->atom_counts.upload(P.atom_counts));
  HIP_TRY(ag->q.upload(P.ap.q));
  HIP_TRY(ag->ljp1.upload(P.ap.ljp1));
  HIP_TRY(ag->ljp2.upload(P.ap.ljp2));
  HIP_TRY(ag->lj_type.upload(P.ap.lj_type));
  HIP_TRY(ag->ljab.upload(P.ts.ljab));
  HIP_TRY(ag->ljab14.upload(P.ts.ljab14));
  HIP_TRY(ag->rho.upload(P.ap.rho));
  HIP_TRY(ag->scr.upload(P.ap.scr));
  HIP_TRY(ag->pbr.upload(P.ap.pbr));
  HIP_TRY(ag->alpha.upload(P.ap.alpha));
  HIP_TRY(ag->beta.upload(P.ap.beta));
  HIP_TRY(ag->gamma.upload(P.ap.gamma));
  HIP_TRY(ag->wu.upload(P.wu));
  HIP_TRY(ag->wu_sys.upload(P.wu_sys));
  HIP_TRY(ag->tasks.upload(P.ts.tasks));
  HIP_TRY(ag->masks.upload(P.ts.masks));
  HIP_TRY(ag->vterms.upload(P.ts.vterms));
  HIP_TRY(ag->bond_par.upload(P.ts.bond_par));
  HIP_TRY(ag->angl_par.upload(P.ts.angl_par));
  HIP_TRY(ag->dihe_par.upload(P.ts.dihe_par));
  HIP_TRY(ag->dihe14.upload(P.ts.dihe14));
  HIP_TRY(ag->rst.upload(P.rst));
  std::vector<int> zero_status(1, 0);
  HIP_TRY(ag->status.upload(zero_status));
  std::vector<int> zi(P.nsys, 0);
  std::vector<double> zd(P.nsys, 0.0);
  HIP_TRY(ag->cycles_run.upload(zi));
  HIP_TRY(ag->rms_force.upload(zd));
  HIP_TRY(ag->stop_flags.upload(zi));
  MMKit &k = ag->kit;
  std::memset(&k, 0, sizeof(k));
  k.nsys = P.nsys;
  k.nwu = static_cast<int>(P.wu.size());
  k.maxa = round_up(std::max(P.maxa, kTile), kTile);
  k.block = P.block;
  k.igb = P.sm.igb;
  k.dielectric = P.sm.dielectric;
  k.kappa = P.sm.kappa;
  k.sys_start = ag->atom_starts.as<int>();
  k.sys_natom = ag->atom_counts.as<int>();
  k.wu = ag->wu.as<WuDesc>();
  k.wu_sys = ag->wu_sys.as<WuSys>();
  k.tasks = ag->tasks.as<Task>();
  k.masks = ag->masks.as<unsigned>();
  k.vterms = ag->vterms.as<int4>();
  k.bond_par = ag->bond_par.as<double2>();
  k.angl_par = ag->angl_par.as<double2>();
  k.dihe_par = ag->dihe_par.as<DiheParam>();
  k.dihe14 = ag->dihe14.as<double2>();
  k.q = ag->q.as<double>();
  k.ljp1 = ag->ljp1.as<double>();
  k.ljp2 = ag->ljp2.as<double>();
  k.lj_type = ag->lj_type.as<int>();
  k.ljab = ag->ljab.as<double2>();
  k.ljab14 = ag->ljab14.as<double2>();
  k.rho = ag->rho.as<double>();
  k.scr = ag->scr.as<double>();
  k.pbr = ag->pbr.as<double>();
  k.gb_alpha = ag->alpha.as<double>();
  k.gb_beta = ag->beta.as<double>();
  k.gb_gamma = ag->gamma.as<double>();
  k.status = ag->status.as<int>();
  k.rst = ag->rst.as<stormm_restraint>();
  *out = ag.release();
  return STORMM_OK;
}

extern "C" int stormm_synthesis_destroy(stormm_synthesis *ag) {
  delete ag;
  return STORMM_OK;
}

extern "C" int stormm_synthesis_get_info(const stormm_synthesis *ag, stormm_synthesis_info *info) {
  if (ag == nullptr || info == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  const SynthesisPlan &P = ag->plan;
  info->system_count = P.nsys;
  info->padded_atom_count = P.padded_atoms;
  info->work_unit_count = static_cast<int>(P.wu.size());
  info->task_count = static_cast<int>(P.task_count);
  info->max_wu_atoms = P.maxa;
  info->lj_table_systems = P.lj_table_systems;
  info->pair_count = P.pairs;
  info->valence_term_count = P.vterm_count;
  return STORMM_OK;
}

extern "C" int stormm_synthesis_atom_layout(const stormm_synthesis *ag, int32_t *atom_starts,
                                            int32_t *atom_counts) {
  if (ag == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null synthesis");
  }
  const SynthesisPlan &P = ag->plan;
  for (int s = 0; s < P.nsys; s++) {
    if (atom_starts) {
      atom_starts[s] = P.atom_starts[s];
    }
    if (atom_counts) {
      atom_counts[s] = P.atom_counts[s];
    }
  }
  return STORMM_OK;
}

// ---- PhaseSpaceSynthesis -----------------------------------------------------------------
// Fixed-precision values cross between host and device as 128-bit integers: the device holds a
// 64-bit word per value and, above 38 coordinate bits or 40 force bits, a 32-bit overflow word
// (value = ovrf 2^63 + word).
static bool wide_coordinates(const stormm_phasespace *ps) {
  return ps->gpos_bits > kMaxPlainGposBits;
}
static bool wide_forces(const stormm_phasespace *ps) { return ps->frc_bits > kMaxPlainFrcBits; }

// hostDoubleToInt95 (src/Numerics/split_fixed_precision.cpp): a value already scaled to the
// fixed-precision grid, rounded half away from zero.
static __int128 host_fixed(double scaled) {
  if (std::fabs(scaled) >= kTwo63) {
    const int spill = static_cast<int>(scaled / kTwo63);
    return i95_value(std::llround(scaled - static_cast<double>(spill) * kTwo63), spill);
  }
  return static_cast<__int128>(std::llround(scaled));
}

extern "C" int stormm_phasespace_create(const stormm_synthesis *ag, int32_t gpos_bits,
                                        int32_t frc_bits, stormm_phasespace **out) {
  if (ag == nullptr || out == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  if (gpos_bits < 24 || gpos_bits > 72) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "gpos_bits must lie in [24, 72]");
  }
  if (frc_bits < 18 || frc_bits > 72) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "frc_bits must lie in [18, 72]");
  }
  std::unique_ptr<stormm_phasespace> ps(new stormm_phasespace());
  ps->ag = ag;
  ps->gpos_bits = gpos_bits;
  ps->frc_bits = frc_bits;
  const size_t na = static_cast<size_t>(ag->plan.padded_atoms);
  DevMem *words[12] = {&ps->xcrd, &ps->ycrd, &ps->zcrd, &ps->xfrc, &ps->yfrc, &ps->zfrc,
                       &ps->xprv, &ps->yprv, &ps->zprv, &ps->xcg, &ps->ycg, &ps->zcg};
  for (DevMem *d : words) {
    HIP_TRY(d->alloc(na * sizeof(long long)));
    HIP_TRY(hipMemset(d->ptr, 0, na * sizeof(long long)));
  }
  DevMem *ovrf[12] = {&ps->xcrd_o, &ps->ycrd_o, &ps->zcrd_o, &ps->xfrc_o, &ps->yfrc_o,
                      &ps->zfrc_o, &ps->xprv_o, &ps->yprv_o, &ps->zprv_o, &ps->xcg_o,
                      &ps->ycg_o, &ps->zcg_o};
  for (DevMem *d : ovrf) {
    HIP_TRY(d->alloc(na * sizeof(int)));
    HIP_TRY(hipMemset(d->ptr, 0, na * sizeof(int)));
  }
  *out = ps.release();
  return STORMM_OK;
}

extern "C" int stormm_phasespace_destroy(stormm_phasespace *ps) {
  delete ps;
  return STORMM_OK;
}

// Systems [lo, hi) occupy one contiguous span of device atoms, system after system with each
// system padded to the wave width; the helpers below move that span in one copy per array.
static int system_span(const stormm_phasespace *ps, int lo, int hi, int *begin, int *end) {
  if (ps == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  const SynthesisPlan &P = ps->ag->plan;
  if (lo < 0 || hi > P.nsys || lo >= hi) {
    return fail(STORMM_ERR_INVALID_ARGUMENT,
                "system range [" + std::to_string(lo) + ", " + std::to_string(hi) +
                ") is empty or outside the synthesis's " + std::to_string(P.nsys) + " systems");
  }
  *begin = P.atom_starts[lo];
  *end = P.atom_starts[hi - 1] + P.atom_counts[hi - 1];
  return STORMM_OK;
}

// Coordinates (forces == false) or forces of device atoms [b, e), per axis.
static int read_span(const stormm_phasespace *ps, bool forces, int b, int e,
                     std::vector<__int128> out[3]) {
  const size_t n = static_cast<size_t>(e - b);
  const DevMem *w[3] = {forces ? &ps->xfrc : &ps->xcrd, forces ? &ps->yfrc : &ps->ycrd,
                        forces ? &ps->zfrc : &ps->zcrd};
  const DevMem *o[3] = {forces ? &ps->xfrc_o : &ps->xcrd_o, forces ? &ps->yfrc_o : &ps->ycrd_o,
                        forces ? &ps->zfrc_o : &ps->zcrd_o};
  const bool wide = forces ? wide_forces(ps) : wide_coordinates(ps);
  std::vector<long long> lo(n);
  std::vector<int> hi(n, 0);
  for (int d = 0; d < 3; d++) {
    HIP_TRY(hipMemcpy(lo.data(), w[d]->as<long long>() + b, n * sizeof(long long),
                      hipMemcpyDeviceToHost));
    if (wide) {
      HIP_TRY(hipMemcpy(hi.data(), o[d]->as<int>() + b, n * sizeof(int), hipMemcpyDeviceToHost));
    }
    out[d].resize(n);
    for (size_t i = 0; i < n; i++) {
      out[d][i] = i95_value(lo[i], hi[i]);
    }
  }
  return STORMM_OK;
}

// Coordinates of device atoms [b, e), per axis (padding atoms keep the values given).
static int write_span_coordinates(stormm_phasespace *ps, int b, int e,
                                  const std::vector<__int128> v[3]) {
  const size_t n = static_cast<size_t>(e - b);
  DevMem *w[3] = {&ps->xcrd, &ps->ycrd, &ps->zcrd};
  DevMem *o[3] = {&ps->xcrd_o, &ps->ycrd_o, &ps->zcrd_o};
  const bool wide = wide_coordinates(ps);
  std::vector<long long> lo(n);
  std::vector<int> hi(n);
  for (int d = 0; d < 3; d++) {
    for (size_t i = 0; i < n; i++) {
      if (wide) {
        i95_split(v[d][i], &lo[i], &hi[i]);
      }
      else {
        const __int128 x = v[d][i];
        if (x > static_cast<__int128>(LLONG_MAX) || x < static_cast<__int128>(LLONG_MIN)) {
          return fail(STORMM_ERR_INVALID_ARGUMENT, "coordinate beyond the 64-bit range of " +
                                                       std::to_string(ps->gpos_bits) + " bits");
        }
        lo[i] = static_cast<long long>(x);
      }
    }
    HIP_TRY(hipMemcpy(w[d]->as<long long>() + b, lo.data(), n * sizeof(long long),
                      hipMemcpyHostToDevice));
    if (wide) {
      HIP_TRY(hipMemcpy(o[d]->as<int>() + b, hi.data(), n * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  return STORMM_OK;
}

// Coordinates of systems [lo, hi) from [atoms][3] fixed-precision values.
static int put_systems(stormm_phasespace *ps, int lo, int hi,
                       const std::function<__int128(size_t)> &value) {
  int b = 0, e = 0;
  int rc = system_span(ps, lo, hi, &b, &e);
  if (rc != STORMM_OK) {
    return rc;
  }
  const SynthesisPlan &P = ps->ag->plan;
  std::vector<__int128> v[3];
  for (int d = 0; d < 3; d++) {
    v[d].assign(static_cast<size_t>(e - b), 0);
  }
  size_t c = 0;
  for (int s = lo; s < hi; s++) {
    for (int i = 0; i < P.atom_counts[s]; i++) {
      const int g = P.atom_starts[s] + i - b;
      for (int d = 0; d < 3; d++) {
        v[d][g] = value(3 * c + d);
      }
      c++;
    }
  }
  return write_span_coordinates(ps, b, e, v);
}

// Coordinates or forces of systems [lo, hi) into [atoms][3] outputs.
static int get_systems(const stormm_phasespace *ps, bool forces, int lo, int hi,
                       const std::function<void(size_t, __int128)> &take) {
  int b = 0, e = 0;
  int rc = system_span(ps, lo, hi, &b, &e);
  if (rc != STORMM_OK) {
    return rc;
  }
  const SynthesisPlan &P = ps->ag->plan;
  std::vector<__int128> v[3];
  if ((rc = read_span(ps, forces, b, e, v)) != STORMM_OK) {
    return rc;
  }
  size_t c = 0;
  for (int s = lo; s < hi; s++) {
    for (int i = 0; i < P.atom_counts[s]; i++) {
      const int g = P.atom_starts[s] + i - b;
      for (int d = 0; d < 3; d++) {
        take(3 * c + d, v[d][g]);
      }
      c++;
    }
  }
  return STORMM_OK;
}

// Real value of a fixed-precision integer in units of 2^-bits.
static double fixed_to_real(__int128 v, int bits) {
  const double inv = std::ldexp(1.0, -bits);
  if (v <= static_cast<__int128>(LLONG_MAX) && v >= static_cast<__int128>(LLONG_MIN)) {
    return static_cast<double>(static_cast<long long>(v)) * inv;
  }
  return i128_to_double(v) * inv;
}

extern "C" int stormm_phasespace_upload_fixed_split(stormm_phasespace *ps, const int64_t *xyz,
                                                    const int32_t *xyz_ovrf) {
  if (ps == nullptr || xyz == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  return put_systems(ps, 0, ps->ag->plan.nsys, [&](size_t k) {
    return i95_value(xyz[k], (xyz_ovrf != nullptr) ? xyz_ovrf[k] : 0);
  });
}

extern "C" int stormm_phasespace_upload_fixed(stormm_phasespace *ps, const int64_t *xyz) {
  return stormm_phasespace_upload_fixed_split(ps, xyz, nullptr);
}

extern "C" int stormm_phasespace_upload_coordinates(stormm_phasespace *ps, const double *xyz) {
  if (ps == nullptr || xyz == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  const double scale = std::ldexp(1.0, ps->gpos_bits);
  return put_systems(ps, 0, ps->ag->plan.nsys,
                     [&](size_t k) { return host_fixed(xyz[k] * scale); });
}

extern "C" int stormm_phasespace_download_fixed_split(const stormm_phasespace *ps, int64_t *xyz,
                                                      int32_t *xyz_ovrf, int64_t *fxyz,
                                                      int32_t *fxyz_ovrf) {
  if (ps == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  for (int pass = 0; pass < 2; pass++) {
    int64_t *w = (pass == 0) ? xyz : fxyz;
    int32_t *o = (pass == 0) ? xyz_ovrf : fxyz_ovrf;
    if (w == nullptr) {
      continue;
    }
    bool overflow = false;
    const int rc = get_systems(ps, pass == 1, 0, ps->ag->plan.nsys, [&](size_t k, __int128 v) {
      if (o != nullptr) {
        long long x;
        int y;
        i95_split(v, &x, &y);
        if (y == 0 || v > static_cast<__int128>(LLONG_MAX) ||
            v < static_cast<__int128>(LLONG_MIN)) {
          w[k] = x;
          o[k] = y;
        }
        else {
          w[k] = static_cast<long long>(v);  // values of one word keep an overflow word of 0
          o[k] = 0;
        }
      }
      else {
        overflow = overflow || v > static_cast<__int128>(LLONG_MAX) ||
                   v < static_cast<__int128>(LLONG_MIN);
        w[k] = static_cast<long long>(v);
      }
    });
    if (rc != STORMM_OK) {
      return rc;
    }
    if (overflow) {
      return fail(STORMM_ERR_STATE, "values exceed 64 bits: download them with their overflow "
                                    "words");
    }
  }
  return STORMM_OK;
}

extern "C" int stormm_phasespace_download_fixed(const stormm_phasespace *ps, int64_t *xyz,
                                                int64_t *fxyz) {
  return stormm_phasespace_download_fixed_split(ps, xyz, nullptr, fxyz, nullptr);
}

extern "C" int stormm_phasespace_download_coordinates(const stormm_phasespace *ps, double *xyz) {
  if (ps == nullptr || xyz == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  return get_systems(ps, false, 0, ps->ag->plan.nsys, [&](size_t k, __int128 v) {
    xyz[k] = fixed_to_real(v, ps->gpos_bits);
  });
}

extern "C" int stormm_phasespace_download_forces(const stormm_phasespace *ps, double *fxyz) {
  if (ps == nullptr || fxyz == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  return get_systems(ps, true, 0, ps->ag->plan.nsys, [&](size_t k, __int128 v) {
    fxyz[k] = fixed_to_real(v, ps->frc_bits);
  });
}

extern "C" int stormm_phasespace_upload_systems(stormm_phasespace *ps, int32_t lo, int32_t hi,
                                                const double *xyz) {
  if (ps == nullptr || xyz == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null phase space or coordinates");
  }
  const double scale = std::ldexp(1.0, ps->gpos_bits);
  return put_systems(ps, lo, hi, [&](size_t k) { return host_fixed(xyz[k] * scale); });
}

extern "C" int stormm_phasespace_download_systems(const stormm_phasespace *ps, int32_t lo,
                                                  int32_t hi, double *xyz, double *fxyz) {
  if (ps == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  for (int pass = 0; pass < 2; pass++) {
    double *dst = (pass == 0) ? xyz : fxyz;
    if (dst == nullptr) {
      continue;
    }
    const int bits = (pass == 0) ? ps->gpos_bits : ps->frc_bits;
    const int rc = get_systems(ps, pass == 1, lo, hi,
                               [&](size_t k, __int128 v) { dst[k] = fixed_to_real(v, bits); });
    if (rc != STORMM_OK) {
      return rc;
    }
  }
  return STORMM_OK;
}

extern "C" int stormm_phasespace_import_system(stormm_phasespace *ps, int32_t system_index,
                                               const double *xyz) {
  return stormm_phasespace_upload_systems(ps, system_index, system_index + 1, xyz);
}

extern "C" int stormm_phasespace_export_system(const stormm_phasespace *ps, int32_t system_index,
                                               double *xyz, double *fxyz) {
  return stormm_phasespace_download_systems(ps, system_index, system_index + 1, xyz, fxyz);
}

extern "C" int stormm_phasespace_writer(stormm_phasespace *ps, stormm_ps_writer *w) {
  if (ps == nullptr || w == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  std::memset(w, 0, sizeof(*w));
  w->system_count = ps->ag->plan.nsys;
  w->gpos_bits = ps->gpos_bits;
  w->frc_bits = ps->frc_bits;
  w->gpos_scale = std::ldexp(1.0, ps->gpos_bits);
  w->inv_gpos_scale = std::ldexp(1.0, -ps->gpos_bits);
  w->frc_scale = std::ldexp(1.0, ps->frc_bits);
  w->inv_frc_scale = std::ldexp(1.0, -ps->frc_bits);
  w->atom_starts = ps->ag->atom_starts.as<int32_t>();
  w->atom_counts = ps->ag->atom_counts.as<int32_t>();
  w->xcrd = ps->xcrd.as<int64_t>();
  w->ycrd = ps->ycrd.as<int64_t>();
  w->zcrd = ps->zcrd.as<int64_t>();
  w->xfrc = ps->xfrc.as<int64_t>();
  w->yfrc = ps->yfrc.as<int64_t>();
  w->zfrc = ps->zfrc.as<int64_t>();
  w->xprv = ps->xprv.as<int64_t>();
  w->yprv = ps->yprv.as<int64_t>();
  w->zprv = ps->zprv.as<int64_t>();
  w->x_cg_temp = ps->xcg.as<int64_t>();
  w->y_cg_temp = ps->ycg.as<int64_t>();
  w->z_cg_temp = ps->zcg.as<int64_t>();
  w->xcrd_ovrf = ps->xcrd_o.as<int32_t>();
  w->ycrd_ovrf = ps->ycrd_o.as<int32_t>();
  w->zcrd_ovrf = ps->zcrd_o.as<int32_t>();
  w->xfrc_ovrf = ps->xfrc_o.as<int32_t>();
  w->yfrc_ovrf = ps->yfrc_o.as<int32_t>();
  w->zfrc_ovrf = ps->zfrc_o.as<int32_t>();
  w->xprv_ovrf = ps->xprv_o.as<int32_t>();
  w->yprv_ovrf = ps->yprv_o.as<int32_t>();
  w->zprv_ovrf = ps->zprv_o.as<int32_t>();
  w->x_cg_temp_ovrf = ps->xcg_o.as<int32_t>();
  w->y_cg_temp_ovrf = ps->ycg_o.as<int32_t>();
  w->z_cg_temp_ovrf = ps->zcg_o.as<int32_t>();
  return STORMM_OK;
}

// Zeroes the words (and, above the plain limits, the overflow words) of up to three groups of
// per-atom arrays of a writer.
static int zero_writer_arrays(int64_t *a, int64_t *b, int64_t *c, int32_t *ao, int32_t *bo,
                              int32_t *co, bool wide, size_t n, hipStream_t st) {
  HIP_TRY(launch_fill(reinterpret_cast<long long *>(a), reinterpret_cast<long long *>(b),
                      reinterpret_cast<long long *>(c), n, st));
  if (wide) {
    if (ao == nullptr || bo == nullptr || co == nullptr) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "overflow arrays are required above 40 force "
                                               "bits");
    }
    HIP_TRY(launch_fill_int(ao, bo, co, n, st));
  }
  return STORMM_OK;
}

extern "C" int stormm_ps_writer_initialize_forces(const stormm_ps_writer *w, int32_t padded_atoms,
                                                  void *stream) {
  if (w == nullptr || padded_atoms < 0 || w->xfrc == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid writer");
  }
  return zero_writer_arrays(w->xfrc, w->yfrc, w->zfrc, w->xfrc_ovrf, w->yfrc_ovrf, w->zfrc_ovrf,
                            w->frc_bits > kMaxPlainFrcBits, static_cast<size_t>(padded_atoms),
                            static_cast<hipStream_t>(stream));
}

extern "C" int stormm_ps_writer_prime_conjugate_gradient(const stormm_ps_writer *w,
                                                         int32_t padded_atoms, void *stream) {
  if (w == nullptr || padded_atoms < 0 || w->xprv == nullptr || w->x_cg_temp == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid writer");
  }
  const bool wide = w->frc_bits > kMaxPlainFrcBits;
  const size_t n = static_cast<size_t>(padded_atoms);
  hipStream_t st = static_cast<hipStream_t>(stream);
  int rc = zero_writer_arrays(w->xprv, w->yprv, w->zprv, w->xprv_ovrf, w->yprv_ovrf,
                              w->zprv_ovrf, wide, n, st);
  if (rc == STORMM_OK) {
    rc = zero_writer_arrays(w->x_cg_temp, w->y_cg_temp, w->z_cg_temp, w->x_cg_temp_ovrf,
                            w->y_cg_temp_ovrf, w->z_cg_temp_ovrf, wide, n, st);
  }
  return rc;
}

extern "C" int stormm_phasespace_initialize_forces(stormm_phasespace *ps, void *stream) {
  if (ps == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  stormm_ps_writer w;
  stormm_phasespace_writer(ps, &w);
  return stormm_ps_writer_initialize_forces(&w, ps->ag->plan.padded_atoms, stream);
}

extern "C" int stormm_phasespace_prime_conjugate_gradient(stormm_phasespace *ps, void *stream) {
  if (ps == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  stormm_ps_writer w;
  stormm_phasespace_writer(ps, &w);
  return stormm_ps_writer_prime_conjugate_gradient(&w, ps->ag->plan.padded_atoms, stream);
}

// ---- ScoreCard -----------------------------------------------------------------------------
extern "C" int stormm_scorecard_create(int32_t system_count, int32_t nrg_bits,
                                       stormm_scorecard **out) {
  if (out == nullptr || system_count <= 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid ScoreCard request");
  }
  if (nrg_bits < 22 || nrg_bits > 40) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "nrg_bits must lie in [22, 40]");
  }
  std::unique_ptr<stormm_scorecard> sc(new stormm_scorecard());
  sc->nsys = system_count;
  sc->nrg_bits = nrg_bits;
  const size_t n = static_cast<size_t>(system_count) * kScoreStride;
  HIP_TRY(sc->inst.alloc(n * sizeof(long long)));
  HIP_TRY(sc->running.alloc(n * sizeof(double)));
  HIP_TRY(sc->squared.alloc(n * sizeof(double)));
  HIP_TRY(hipMemset(sc->inst.ptr, 0, n * sizeof(long long)));
  HIP_TRY(hipMemset(sc->running.ptr, 0, n * sizeof(double)));
  HIP_TRY(hipMemset(sc->squared.ptr, 0, n * sizeof(double)));
  *out = sc.release();
  return STORMM_OK;
}

extern "C" int stormm_scorecard_destroy(stormm_scorecard *sc) {
  delete sc;
  return STORMM_OK;
}

extern "C" int stormm_scorecard_initialize(stormm_scorecard *sc, void *stream) {
  if (sc == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  HIP_TRY(launch_fill(sc->inst.as<long long>(), nullptr, nullptr,
                      static_cast<size_t>(sc->nsys) * kScoreStride,
                      static_cast<hipStream_t>(stream)));
  return STORMM_OK;
}

// Sizes the time series for at least `need` samples and allocates it if it is not yet; the
// samples held are kept (their copy waits for the device to finish its queued work).
static int scorecard_series(stormm_scorecard *sc, int need) {
  if (sc->series.ptr != nullptr && need <= sc->capacity) {
    return STORMM_OK;
  }
  const int cap = std::max(need, sc->capacity);
  const size_t row = static_cast<size_t>(sc->nsys) * kScoreStride * sizeof(long long);
  DevMem fresh;
  HIP_TRY(fresh.alloc(row * static_cast<size_t>(cap)));
  HIP_TRY(hipMemset(fresh.ptr, 0, fresh.bytes));
  if (sc->series.ptr != nullptr && sc->samples > 0) {
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(fresh.ptr, sc->series.ptr, row * static_cast<size_t>(sc->samples),
                      hipMemcpyDeviceToDevice));
  }
  std::swap(fresh.ptr, sc->series.ptr);
  std::swap(fresh.bytes, sc->series.bytes);
  sc->capacity = cap;
  return STORMM_OK;
}

extern "C" int stormm_scorecard_commit(stormm_scorecard *sc, void *stream) {
  if (sc == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  // A full series doubles, as ScoreCard::commit does (src/Potential/scorecard.cpp:368-372).
  int rc = scorecard_series(sc, (sc->samples < sc->capacity) ? sc->capacity : 2 * sc->capacity);
  if (rc != STORMM_OK) {
    return rc;
  }
  long long *row = sc->series.as<long long>() +
                   static_cast<size_t>(sc->samples) * sc->nsys * kScoreStride;
  HIP_TRY(launch_scorecard_commit(sc->inst.as<long long>(), sc->running.as<double>(),
                                  sc->squared.as<double>(), row, sc->nsys,
                                  std::ldexp(1.0, -sc->nrg_bits), st));
  sc->samples += 1;
  sc->time_steps.push_back(0);
  return STORMM_OK;
}

extern "C" int stormm_scorecard_reserve(stormm_scorecard *sc, int32_t capacity) {
  if (sc == nullptr || capacity < 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid reserve request");
  }
  return scorecard_series(sc, std::max(capacity, sc->samples));
}

extern "C" int stormm_scorecard_reset_sample_count(stormm_scorecard *sc, int32_t count) {
  if (sc == nullptr || count < 0 || count > sc->samples) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "the sample count can be reset to a value between 0 "
                                             "and the samples held");
  }
  sc->samples = count;
  sc->time_steps.resize(static_cast<size_t>(count));
  return STORMM_OK;
}

extern "C" int stormm_scorecard_set_last_time_step(stormm_scorecard *sc, int32_t time_step) {
  if (sc == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  if (sc->samples <= 0) {
    return fail(STORMM_ERR_STATE, "no energies have been committed to the history");
  }
  sc->time_steps[static_cast<size_t>(sc->samples) - 1] = time_step;
  return STORMM_OK;
}

extern "C" int stormm_scorecard_sample_count(const stormm_scorecard *sc, int32_t *samples,
                                             int32_t *capacity) {
  if (sc == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  if (samples != nullptr) {
    *samples = sc->samples;
  }
  if (capacity != nullptr) {
    *capacity = sc->capacity;
  }
  return STORMM_OK;
}

extern "C" int stormm_scorecard_download_series(const stormm_scorecard *sc, int32_t first,
                                                int32_t count, int64_t *series,
                                                int32_t *time_steps) {
  if (sc == nullptr || first < 0 || count < 0 || first + count > sc->samples ||
      (count > 0 && series == nullptr)) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "the series request must lie within the samples "
                                             "held");
  }
  if (count == 0) {
    return STORMM_OK;
  }
  const size_t row = static_cast<size_t>(sc->nsys) * kScoreStride;
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(series, sc->series.as<long long>() + row * first,
                    row * count * sizeof(long long), hipMemcpyDeviceToHost));
  if (time_steps != nullptr) {
    for (int i = 0; i < count; i++) {
      time_steps[i] = sc->time_steps[static_cast<size_t>(first + i)];
    }
  }
  return STORMM_OK;
}

extern "C" int stormm_scorecard_download(const stormm_scorecard *sc, int64_t *acc) {
  if (sc == nullptr || acc == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  HIP_TRY(hipMemcpy(acc, sc->inst.ptr, static_cast<size_t>(sc->nsys) * kScoreStride *
                                           sizeof(long long), hipMemcpyDeviceToHost));
  return STORMM_OK;
}

extern "C" int stormm_scorecard_download_history(const stormm_scorecard *sc, double *running,
                                                 double *squared, int32_t *sample_count) {
  if (sc == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  const size_t n = static_cast<size_t>(sc->nsys) * kScoreStride * sizeof(double);
  if (running) {
    HIP_TRY(hipMemcpy(running, sc->running.ptr, n, hipMemcpyDeviceToHost));
  }
  if (squared) {
    HIP_TRY(hipMemcpy(squared, sc->squared.ptr, n, hipMemcpyDeviceToHost));
  }
  if (sample_count) {
    *sample_count = sc->samples;
  }
  return STORMM_OK;
}

extern "C" int64_t *stormm_scorecard_device_accumulators(stormm_scorecard *sc) {
  return (sc == nullptr) ? nullptr : sc->inst.as<int64_t>();
}

extern "C" int stormm_scorecard_data(stormm_scorecard *sc, stormm_scorecard_writer *w) {
  if (sc == nullptr || w == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  int rc = scorecard_series(sc, sc->capacity);
  if (rc != STORMM_OK) {
    return rc;
  }
  w->system_count = sc->nsys;
  w->data_stride = kScoreStride;
  w->nrg_bits = sc->nrg_bits;
  w->instantaneous_accumulators = sc->inst.as<int64_t>();
  w->sampled_step_count = sc->samples;
  w->sample_capacity = sc->capacity;
  w->running_accumulators = sc->running.as<double>();
  w->squared_accumulators = sc->squared.as<double>();
  w->time_series = sc->series.as<int64_t>();
  return STORMM_OK;
}

extern "C" int stormm_scorecard_writer_initialize(const stormm_scorecard_writer *w,
                                                  void *stream) {
  if (w == nullptr || w->instantaneous_accumulators == nullptr || w->system_count < 0 ||
      w->data_stride <= 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid ScoreCard writer");
  }
  HIP_TRY(launch_fill(reinterpret_cast<long long *>(w->instantaneous_accumulators), nullptr,
                      nullptr, static_cast<size_t>(w->system_count) * w->data_stride,
                      static_cast<hipStream_t>(stream)));
  return STORMM_OK;
}

// ---- Controls and line-minimization records ------------------------------------------------
extern "C" int stormm_mm_controls_create(const stormm_minimize_controls *mincon,
                                         stormm_mm_controls **out) {
  if (out == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null output");
  }
  stormm_mm_controls *c = new stormm_mm_controls();
  if (mincon) {
    c->mc = *mincon;
  }
  else {
    c->mc.total_cycles = 200;
    c->mc.steepest_descent_cycles = 50;
    c->mc.initial_step = 0.01;
    c->mc.convergence_rms_force = 0.0;
    c->mc.clash_damping_cycles = 25;
    c->mc.clash_minimum_distance = 0.2;
    c->mc.clash_vdw_ratio = 0.5;
  }
  *out = c;
  return STORMM_OK;
}

extern "C" int stormm_mm_controls_destroy(stormm_mm_controls *ctrl) {
  delete ctrl;
  return STORMM_OK;
}

extern "C" int stormm_mm_controls_get_step(const stormm_mm_controls *ctrl, int32_t *step) {
  if (ctrl == nullptr || step == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  *step = ctrl->step;
  return STORMM_OK;
}

extern "C" int stormm_mm_controls_set_step(stormm_mm_controls *ctrl, int32_t step) {
  if (ctrl == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  ctrl->step = step;
  return STORMM_OK;
}

extern "C" int stormm_mm_controls_data(const stormm_mm_controls *ctrl, stormm_mm_kit *kit) {
  if (ctrl == nullptr || kit == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  kit->step = ctrl->step;
  kit->sd_cycles = ctrl->mc.steepest_descent_cycles;
  return STORMM_OK;
}

extern "C" int stormm_line_min_create(int32_t system_count, double dx0, stormm_line_min **out) {
  if (out == nullptr || system_count <= 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid line-minimization request");
  }
  std::unique_ptr<stormm_line_min> lm(new stormm_line_min());
  lm->nsys = system_count;
  std::vector<double> v(static_cast<size_t>(9) * system_count, 0.0);
  for (int s = 0; s < system_count; s++) {
    v[s] = dx0;  // l_move, LineMinimization::primeMoveLengths (line_minimization.cpp:340)
  }
  HIP_TRY(lm->vals.upload(v));
  *out = lm.release();
  return STORMM_OK;
}

extern "C" int stormm_line_min_destroy(stormm_line_min *lm) {
  delete lm;
  return STORMM_OK;
}

extern "C" int stormm_line_min_data(stormm_line_min *lm, stormm_linmin_writer *w) {
  if (lm == nullptr || w == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  double *v = lm->vals.as<double>();
  const size_t n = static_cast<size_t>(lm->nsys);
  w->system_count = lm->nsys;
  w->l_move = v;
  w->s_move = v + n;
  w->mfac_a = v + 2 * n;
  w->mfac_b = v + 3 * n;
  w->mfac_c = v + 4 * n;
  w->nrg_a = v + 5 * n;
  w->nrg_b = v + 6 * n;
  w->nrg_c = v + 7 * n;
  w->nrg_d = v + 8 * n;
  return STORMM_OK;
}

extern "C" int stormm_line_min_download(const stormm_line_min *lm, double *values9xN) {
  if (lm == nullptr || values9xN == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  HIP_TRY(hipMemcpy(values9xN, lm->vals.ptr, static_cast<size_t>(9) * lm->nsys * sizeof(double),
                    hipMemcpyDeviceToHost));
  return STORMM_OK;
}

// ---- Launches ------------------------------------------------------------------------------
// Every launch runs on writer views: the handle entry points take the views of their objects,
// the kit entry points the caller's.
struct ScView {
  int nsys = 0, nrg_bits = 26;
  long long *acc = nullptr;
  // Energy history: the time series, `samples` of its `capacity` samples used, and the running
  // sums (null series: no history).
  long long *series = nullptr;
  double *running = nullptr, *squared = nullptr;
  int samples = 0, capacity = 0;
  stormm_scorecard *owner = nullptr;  // the handle, when the view is of one
  HistKit hist = {};                  // a minimization's energy snapshots (null rows: none)
};
struct LmView {
  int nsys = 0;
  double *v[9] = {};
};

static ScView sc_view(stormm_scorecard *sc) {
  ScView v;
  if (sc != nullptr) {
    v.nsys = sc->nsys;
    v.nrg_bits = sc->nrg_bits;
    v.acc = sc->inst.as<long long>();
    v.series = sc->series.as<long long>();
    v.running = sc->running.as<double>();
    v.squared = sc->squared.as<double>();
    v.samples = sc->samples;
    v.capacity = sc->capacity;
    v.owner = sc;
  }
  return v;
}

static int sc_view(const stormm_scorecard_writer *w, ScView *out) {
  if (w == nullptr) {
    return STORMM_OK;
  }
  if (w->instantaneous_accumulators == nullptr || w->data_stride != kScoreStride ||
      w->nrg_bits < 22 || w->nrg_bits > 40) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "ScoreCard writer needs accumulators, a data "
                                             "stride of 32 and energy bits in [22, 40]");
  }
  if ((w->running_accumulators == nullptr) != (w->squared_accumulators == nullptr) ||
      (w->time_series != nullptr &&
       (w->sampled_step_count < 0 || w->sample_capacity < w->sampled_step_count))) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "ScoreCard writer history needs both running-sum "
                                             "arrays or neither, and 0 <= sampled_step_count "
                                             "<= sample_capacity");
  }
  out->nsys = w->system_count;
  out->nrg_bits = w->nrg_bits;
  out->acc = reinterpret_cast<long long *>(w->instantaneous_accumulators);
  out->series = reinterpret_cast<long long *>(w->time_series);
  out->running = w->running_accumulators;
  out->squared = w->squared_accumulators;
  out->samples = w->sampled_step_count;
  out->capacity = w->sample_capacity;
  return STORMM_OK;
}

// Samples a minimization of mc->total_cycles cycles commits at its diagnostic interval.
static int history_samples(const stormm_minimize_controls *mc) {
  const int every = mc->diagnostic_print_frequency;
  return (every > 0) ? (mc->total_cycles + every - 1) / every : 0;
}

// Readies sc's history for a minimization's snapshots (sc->hist): a handle's series grows as
// needed, a writer's must have room.  No interval or no series: no snapshots.
static int prepare_history(const stormm_minimize_controls *mc, ScView *sc) {
  sc->hist = HistKit{};
  const int count = history_samples(mc);
  if (count == 0) {
    return STORMM_OK;
  }
  if (sc->owner != nullptr) {
    const int rc = scorecard_series(sc->owner, sc->owner->samples + count);
    if (rc != STORMM_OK) {
      return rc;
    }
    *sc = sc_view(sc->owner);
  }
  if (sc->series == nullptr) {
    return STORMM_OK;
  }
  if (sc->samples + count > sc->capacity) {
    return fail(STORMM_ERR_CAPACITY,
                "the ScoreCard's time series holds " + std::to_string(sc->capacity) +
                    " samples; the minimization adds " + std::to_string(count) + " to " +
                    std::to_string(sc->samples));
  }
  HistKit &h = sc->hist;
  h.rows = sc->series;
  h.running = sc->running;
  h.squared = sc->squared;
  h.scorecard = sc->acc;
  h.sample_stride = static_cast<long long>(sc->nsys) * kScoreStride;
  h.sys_stride = kScoreStride;
  h.every = mc->diagnostic_print_frequency;
  h.first = sc->samples;
  h.capacity = sc->capacity;
  h.max_cycle = 0;
  h.inv_nrg_scale = std::ldexp(1.0, -sc->nrg_bits);
  return STORMM_OK;
}

// After the minimization's launches: samples of systems that stopped early repeat their last
// snapshot (steps: per-system cycle counts, null when every system ran every cycle), and a
// handle counts the new samples, labelled with their cycles.
static int finish_history(const stormm_minimize_controls *mc, const ScView &sc,
                          const int *steps, hipStream_t stream) {
  if (sc.hist.rows == nullptr) {
    return STORMM_OK;
  }
  const int count = history_samples(mc);
  if (steps != nullptr && mc->convergence_rms_force > 0.0) {
    HIP_TRY(launch_history_fill(sc.hist, steps, sc.nsys, count, stream));
  }
  if (sc.owner != nullptr) {
    for (int i = 0; i < count; i++) {
      sc.owner->time_steps.push_back(i * mc->diagnostic_print_frequency);
    }
    sc.owner->samples += count;
  }
  return STORMM_OK;
}

static LmView lm_view(stormm_line_min *lm) {
  LmView v;
  v.nsys = lm->nsys;
  for (int j = 0; j < 9; j++) {
    v.v[j] = lm->vals.as<double>() + static_cast<size_t>(j) * lm->nsys;
  }
  return v;
}

static int lm_view(const stormm_linmin_writer *w, LmView *out) {
  double *f[9] = {w->l_move, w->s_move, w->mfac_a, w->mfac_b, w->mfac_c,
                  w->nrg_a,  w->nrg_b,  w->nrg_c,  w->nrg_d};
  out->nsys = w->system_count;
  for (int j = 0; j < 9; j++) {
    if (f[j] == nullptr) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "line-minimization writer with a null array");
    }
    out->v[j] = f[j];
  }
  return STORMM_OK;
}

static int check_objects(const stormm_synthesis *ag, const stormm_phasespace *ps,
                         const stormm_scorecard *sc) {
  if (ag == nullptr || ps == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null synthesis or phase space");
  }
  if (ps->ag != ag) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "phase space was built for another synthesis");
  }
  if (sc != nullptr && sc->nsys != ag->plan.nsys) {
    // src/MolecularMechanics/hpc_minimization.cu:151-155
    return fail(STORMM_ERR_INVALID_ARGUMENT,
                "The energy tracking object is not prepared for the number of systems contained "
                "in the topology synthesis (" + std::to_string(sc->nsys) + " vs. " +
                    std::to_string(ag->plan.nsys) + ").");
  }
  return STORMM_OK;
}

// A writer fit for the synthesis: system count, bit ranges, the arrays a launch touches, and
// the overflow arrays above the plain limits.  `cg`: the conjugate-gradient arrays are needed.
static int check_writer(const stormm_synthesis *ag, const stormm_ps_writer *w, int precision,
                        const ScView *sc, bool cg) {
  if (ag == nullptr || w == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null synthesis or phase-space writer");
  }
  if (precision != STORMM_PREC_SINGLE && precision != STORMM_PREC_DOUBLE) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "precision must be STORMM_PREC_SINGLE or "
                                             "STORMM_PREC_DOUBLE");
  }
  if (w->system_count != ag->plan.nsys) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "the phase-space writer holds " +
                                                 std::to_string(w->system_count) +
                                                 " systems, the synthesis " +
                                                 std::to_string(ag->plan.nsys));
  }
  if (sc != nullptr && sc->acc != nullptr && sc->nsys != ag->plan.nsys) {
    // src/MolecularMechanics/hpc_minimization.cu:151-155
    return fail(STORMM_ERR_INVALID_ARGUMENT,
                "The energy tracking object is not prepared for the number of systems contained "
                "in the topology synthesis (" + std::to_string(sc->nsys) + " vs. " +
                    std::to_string(ag->plan.nsys) + ").");
  }
  if (w->gpos_bits < 24 || w->gpos_bits > 72 || w->frc_bits < 18 || w->frc_bits > 72) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "fixed-precision bits out of range (coordinates "
                                             "24-72, forces 18-72)");
  }
  if (precision == STORMM_PREC_SINGLE && w->frc_bits > kMaxPlainFrcBits) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "single-precision launches take at most 40 force "
                                             "bits");
  }
  if (w->xcrd == nullptr || w->ycrd == nullptr || w->zcrd == nullptr || w->xfrc == nullptr ||
      w->yfrc == nullptr || w->zfrc == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "phase-space writer with null coordinate or force "
                                             "arrays");
  }
  if (w->gpos_bits > kMaxPlainGposBits &&
      (w->xcrd_ovrf == nullptr || w->ycrd_ovrf == nullptr || w->zcrd_ovrf == nullptr)) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "coordinates above 38 bits need overflow arrays");
  }
  if (w->frc_bits > kMaxPlainFrcBits &&
      (w->xfrc_ovrf == nullptr || w->yfrc_ovrf == nullptr || w->zfrc_ovrf == nullptr)) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "forces above 40 bits need overflow arrays");
  }
  if (cg) {
    if (w->xprv == nullptr || w->yprv == nullptr || w->zprv == nullptr ||
        w->x_cg_temp == nullptr || w->y_cg_temp == nullptr || w->z_cg_temp == nullptr) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "conjugate gradient needs the previous-force and "
                                               "direction arrays");
    }
    if (w->frc_bits > kMaxPlainFrcBits &&
        (w->xprv_ovrf == nullptr || w->yprv_ovrf == nullptr || w->zprv_ovrf == nullptr ||
         w->x_cg_temp_ovrf == nullptr || w->y_cg_temp_ovrf == nullptr ||
         w->z_cg_temp_ovrf == nullptr)) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "conjugate gradient above 40 force bits needs "
                                               "the overflow arrays");
    }
  }
  return STORMM_OK;
}

static MMKit make_kit(const stormm_synthesis *ag, const stormm_ps_writer &w, const ScView *sc) {
  MMKit k = ag->kit;
  k.gpos_bits = w.gpos_bits;
  // Forces above 40 bits are accumulated at 40 bits inside a work unit and shifted on store.
  k.frc_bits = std::min(w.frc_bits, kMaxPlainFrcBits);
  k.frc_store_shift = w.frc_bits - k.frc_bits;
  k.nrg_bits = (sc != nullptr && sc->acc != nullptr) ? sc->nrg_bits : 26;
  k.xcrd = reinterpret_cast<long long *>(w.xcrd);
  k.ycrd = reinterpret_cast<long long *>(w.ycrd);
  k.zcrd = reinterpret_cast<long long *>(w.zcrd);
  k.xfrc = reinterpret_cast<long long *>(w.xfrc);
  k.yfrc = reinterpret_cast<long long *>(w.yfrc);
  k.zfrc = reinterpret_cast<long long *>(w.zfrc);
  const bool wide_c = w.gpos_bits > kMaxPlainGposBits;
  const bool wide_f = w.frc_bits > kMaxPlainFrcBits;
  k.xcrd_ovrf = wide_c ? w.xcrd_ovrf : nullptr;
  k.ycrd_ovrf = wide_c ? w.ycrd_ovrf : nullptr;
  k.zcrd_ovrf = wide_c ? w.zcrd_ovrf : nullptr;
  k.xfrc_ovrf = wide_f ? w.xfrc_ovrf : nullptr;
  k.yfrc_ovrf = wide_f ? w.yfrc_ovrf : nullptr;
  k.zfrc_ovrf = wide_f ? w.zfrc_ovrf : nullptr;
  k.scorecard = (sc != nullptr) ? sc->acc : nullptr;
  k.clash_variant = 0;
  k.clash_cycles = 0;
  k.clash_cycle = 0;
  k.clash_step_offset = 0;
  k.clash_step = nullptr;
  k.clash_r0 = 0.0;
  k.clash_ratio = 0.0;
  k.rst_step = 0;
  k.rst_step_sys = nullptr;
  k.rst_cycle_offset = 0;
  k.rst_move = 0;
  k.rst_clash_cycles = 0;
  return k;
}

// Whether the damped kernels run for clash parameters (d, r): the reference launches them when
// either exceeds 10^-6 (src/Potential/hpc_nonbonded_potential.cu:1476, :1662 in single
// precision).
static bool clash_active(int32_t precision, double d, double r) {
  if (precision == STORMM_PREC_DOUBLE) {
    return d >= 1.0e-6 || r >= 1.0e-6;
  }
  return static_cast<float>(d) > 0.000001f || static_cast<float>(r) > 0.000001f;
}

// Kit fields for one evaluation with fixed clash parameters (both zero: plain potentials).
static void set_clash(MMKit &k, int32_t precision, double d, double r) {
  const bool on = clash_active(precision, d, r);
  k.clash_variant = on ? 1 : 0;
  k.clash_cycles = 0;
  k.clash_cycle = 0;
  k.clash_step = nullptr;
  k.clash_step_offset = 0;
  k.clash_r0 = on ? d : 0.0;
  k.clash_ratio = on ? r : 0.0;
}

// Validated minimization controls (src/Namelists/nml_minimize.cpp:227-260, DIE policy).
static int check_mincon(const stormm_minimize_controls *mc) {
  if (mc->total_cycles < 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "negative cycle count");
  }
  if (mc->clash_damping_cycles < 0 || mc->clash_damping_cycles > mc->total_cycles) {
    return fail(STORMM_ERR_INVALID_ARGUMENT,
                "clash damping cycles (" + std::to_string(mc->clash_damping_cycles) +
                    ") must lie between 0 and the total cycle count (" +
                    std::to_string(mc->total_cycles) + ")");
  }
  if (mc->clash_minimum_distance < 0.0 || mc->clash_vdw_ratio < 0.0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "clash parameters must be non-negative");
  }
  if (mc->diagnostic_print_frequency < 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "the diagnostic print frequency must be "
                                             "non-negative");
  }
  return STORMM_OK;
}

static int launch_terms(int32_t precision, const stormm_synthesis *ag,
                        const stormm_ps_writer &w, const ScView &sc, int terms,
                        int32_t eval_force, int32_t eval_energy, double clash_d, double clash_r,
                        int step, void *stream) {
  int rc = check_writer(ag, &w, precision, &sc, false);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (sc.acc == nullptr && eval_energy) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "energy evaluation needs a ScoreCard");
  }
  if (!(clash_d >= 0.0) || !(clash_r >= 0.0)) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "clash parameters must be non-negative");
  }
  MMKit k = make_kit(ag, w, &sc);
  set_clash(k, precision, clash_d, clash_r);
  k.rst_step = step;
  HIP_TRY(launch_mm_eval(precision == STORMM_PREC_DOUBLE ? 0 : 1, k, terms, eval_force ? 1 : 0,
                         eval_energy ? 1 : 0, static_cast<hipStream_t>(stream)));
  return STORMM_OK;
}

static int launch_terms_handles(int32_t precision, const stormm_synthesis *ag,
                                stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                stormm_scorecard *sc, int terms, int32_t eval_force,
                                int32_t eval_energy, double clash_d, double clash_r,
                                void *stream) {
  int rc = check_objects(ag, ps, sc);
  if (rc != STORMM_OK) {
    return rc;
  }
  stormm_ps_writer w;
  stormm_phasespace_writer(ps, &w);
  return launch_terms(precision, ag, w, sc_view(sc), terms, eval_force, eval_energy, clash_d,
                      clash_r, (ctrl != nullptr) ? ctrl->step : 0, stream);
}

static int launch_terms_kit(int32_t precision, const stormm_synthesis *ag,
                            const stormm_mm_kit *ctrl, const stormm_ps_writer *psw,
                            const stormm_scorecard_writer *scw, int terms, int32_t eval_force,
                            int32_t eval_energy, double clash_d, double clash_r, void *stream) {
  ScView sc;
  int rc = sc_view(scw, &sc);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (psw == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null phase-space writer");
  }
  return launch_terms(precision, ag, *psw, sc, terms, eval_force, eval_energy, clash_d, clash_r,
                      (ctrl != nullptr) ? ctrl->step : 0, stream);
}

extern "C" int stormm_launch_nonbonded(int32_t precision, const stormm_synthesis *ag,
                                       stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                       stormm_scorecard *sc, int32_t eval_force,
                                       int32_t eval_energy, double clash_minimum_distance,
                                       double clash_ratio, void *stream) {
  return launch_terms_handles(precision, ag, ctrl, ps, sc, kTermNonbonded, eval_force,
                              eval_energy, clash_minimum_distance, clash_ratio, stream);
}

extern "C" int stormm_launch_valence(int32_t precision, const stormm_synthesis *ag,
                                     stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                     stormm_scorecard *sc, int32_t eval_force,
                                     int32_t eval_energy, double clash_minimum_distance,
                                     double clash_ratio, void *stream) {
  return launch_terms_handles(precision, ag, ctrl, ps, sc, kTermValence, eval_force,
                              eval_energy, clash_minimum_distance, clash_ratio, stream);
}

extern "C" int stormm_launch_mm_evaluation(int32_t precision, const stormm_synthesis *ag,
                                           stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                           stormm_scorecard *sc, int32_t eval_force,
                                           int32_t eval_energy, double clash_minimum_distance,
                                           double clash_ratio, void *stream) {
  return launch_terms_handles(precision, ag, ctrl, ps, sc, kTermNonbonded | kTermValence,
                              eval_force, eval_energy, clash_minimum_distance, clash_ratio,
                              stream);
}

extern "C" int stormm_launch_nonbonded_kit(int32_t precision, const stormm_synthesis *ag,
                                           const stormm_mm_kit *ctrl,
                                           const stormm_ps_writer *psw,
                                           const stormm_scorecard_writer *scw,
                                           int32_t eval_force, int32_t eval_energy,
                                           double clash_minimum_distance, double clash_ratio,
                                           void *stream) {
  return launch_terms_kit(precision, ag, ctrl, psw, scw, kTermNonbonded, eval_force,
                          eval_energy, clash_minimum_distance, clash_ratio, stream);
}

extern "C" int stormm_launch_valence_kit(int32_t precision, const stormm_synthesis *ag,
                                         const stormm_mm_kit *ctrl, const stormm_ps_writer *psw,
                                         const stormm_scorecard_writer *scw, int32_t eval_force,
                                         int32_t eval_energy, double clash_minimum_distance,
                                         double clash_ratio, void *stream) {
  return launch_terms_kit(precision, ag, ctrl, psw, scw, kTermValence, eval_force, eval_energy,
                          clash_minimum_distance, clash_ratio, stream);
}

extern "C" int stormm_launch_mm_evaluation_kit(int32_t precision, const stormm_synthesis *ag,
                                               const stormm_mm_kit *ctrl,
                                               const stormm_ps_writer *psw,
                                               const stormm_scorecard_writer *scw,
                                               int32_t eval_force, int32_t eval_energy,
                                               double clash_minimum_distance, double clash_ratio,
                                               void *stream) {
  return launch_terms_kit(precision, ag, ctrl, psw, scw, kTermNonbonded | kTermValence,
                          eval_force, eval_energy, clash_minimum_distance, clash_ratio, stream);
}

static CgKit make_cg_kit(const stormm_synthesis *ag, const stormm_ps_writer &w, int precision) {
  CgKit k;
  std::memset(&k, 0, sizeof(k));
  k.nsys = ag->plan.nsys;
  k.sys_start = ag->atom_starts.as<int>();
  k.sys_natom = ag->atom_counts.as<int>();
  k.xfrc = reinterpret_cast<long long *>(w.xfrc);
  k.yfrc = reinterpret_cast<long long *>(w.yfrc);
  k.zfrc = reinterpret_cast<long long *>(w.zfrc);
  k.xprv = reinterpret_cast<long long *>(w.xprv);
  k.yprv = reinterpret_cast<long long *>(w.yprv);
  k.zprv = reinterpret_cast<long long *>(w.zprv);
  k.xcg = reinterpret_cast<long long *>(w.x_cg_temp);
  k.ycg = reinterpret_cast<long long *>(w.y_cg_temp);
  k.zcg = reinterpret_cast<long long *>(w.z_cg_temp);
  if (w.frc_bits > kMaxPlainFrcBits) {
    k.xfrc_ovrf = w.xfrc_ovrf;
    k.yfrc_ovrf = w.yfrc_ovrf;
    k.zfrc_ovrf = w.zfrc_ovrf;
    k.xprv_ovrf = w.xprv_ovrf;
    k.yprv_ovrf = w.yprv_ovrf;
    k.zprv_ovrf = w.zprv_ovrf;
    k.xcg_ovrf = w.x_cg_temp_ovrf;
    k.ycg_ovrf = w.y_cg_temp_ovrf;
    k.zcg_ovrf = w.z_cg_temp_ovrf;
  }
  k.inv_frc_scale = std::ldexp(1.0, -w.frc_bits);
  k.dp_mode = (precision == STORMM_PREC_DOUBLE) ? 1 : 0;
  return k;
}

static int conjugate_gradient(int32_t precision, const stormm_synthesis *ag,
                              const stormm_ps_writer &w, int step, int sd_cycles,
                              void *stream) {
  int rc = check_writer(ag, &w, precision, nullptr, true);
  if (rc != STORMM_OK) {
    return rc;
  }
  HIP_TRY(launch_conjugate_gradient(make_cg_kit(ag, w, precision), step, sd_cycles,
                                    static_cast<hipStream_t>(stream)));
  return STORMM_OK;
}

extern "C" int stormm_launch_conjugate_gradient(int32_t precision, const stormm_synthesis *ag,
                                                stormm_phasespace *ps, stormm_mm_controls *ctrl,
                                                void *stream) {
  int rc = check_objects(ag, ps, nullptr);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (ctrl == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null controls");
  }
  stormm_ps_writer w;
  stormm_phasespace_writer(ps, &w);
  return conjugate_gradient(precision, ag, w, ctrl->step, ctrl->mc.steepest_descent_cycles,
                            stream);
}

extern "C" int stormm_launch_conjugate_gradient_kit(int32_t precision, const stormm_synthesis *ag,
                                                    const stormm_ps_writer *psw,
                                                    const stormm_mm_kit *ctrl, void *stream) {
  if (psw == nullptr || ctrl == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null writer or controls kit");
  }
  return conjugate_gradient(precision, ag, *psw, ctrl->step, ctrl->sd_cycles, stream);
}

static LineKit make_line_kit(const stormm_synthesis *ag, const stormm_ps_writer &w,
                             const ScView &sc, const LmView &lm, int precision) {
  LineKit k;
  std::memset(&k, 0, sizeof(k));
  k.nsys = ag->plan.nsys;
  k.sys_start = ag->atom_starts.as<int>();
  k.sys_natom = ag->atom_counts.as<int>();
  k.xcrd = reinterpret_cast<long long *>(w.xcrd);
  k.ycrd = reinterpret_cast<long long *>(w.ycrd);
  k.zcrd = reinterpret_cast<long long *>(w.zcrd);
  k.xfrc = reinterpret_cast<const long long *>(w.xfrc);
  k.yfrc = reinterpret_cast<const long long *>(w.yfrc);
  k.zfrc = reinterpret_cast<const long long *>(w.zfrc);
  if (w.gpos_bits > kMaxPlainGposBits) {
    k.xcrd_ovrf = w.xcrd_ovrf;
    k.ycrd_ovrf = w.ycrd_ovrf;
    k.zcrd_ovrf = w.zcrd_ovrf;
  }
  if (w.frc_bits > kMaxPlainFrcBits) {
    k.xfrc_ovrf = w.xfrc_ovrf;
    k.yfrc_ovrf = w.yfrc_ovrf;
    k.zfrc_ovrf = w.zfrc_ovrf;
  }
  k.scorecard = sc.acc;
  k.inv_nrg_scale = std::ldexp(1.0, -sc.nrg_bits);
  k.inv_frc_scale = std::ldexp(1.0, -w.frc_bits);
  k.gpos_scale = std::ldexp(1.0, w.gpos_bits);
  k.l_move = lm.v[0];
  k.s_move = lm.v[1];
  k.mfac_a = lm.v[2];
  k.mfac_b = lm.v[3];
  k.mfac_c = lm.v[4];
  k.nrg_a = lm.v[5];
  k.nrg_b = lm.v[6];
  k.nrg_c = lm.v[7];
  k.nrg_d = lm.v[8];
  k.dp_mode = (precision == STORMM_PREC_DOUBLE) ? 1 : 0;
  return k;
}

static int line_advance(int32_t precision, const stormm_synthesis *ag, const stormm_ps_writer &w,
                        const ScView &sc, const LmView &lm, int move_number, void *stream) {
  int rc = check_writer(ag, &w, precision, &sc, false);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (sc.acc == nullptr || lm.nsys != ag->plan.nsys) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "line advance needs a ScoreCard and a matching "
                                             "LineMinimization record");
  }
  if (move_number < 0 || move_number > 3) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "move_number must be 0..3");
  }
  HIP_TRY(launch_line_advance(make_line_kit(ag, w, sc, lm, precision), move_number,
                              static_cast<hipStream_t>(stream)));
  return STORMM_OK;
}

extern "C" int stormm_launch_line_advance(int32_t precision, const stormm_synthesis *ag,
                                          stormm_phasespace *ps, stormm_scorecard *sc,
                                          stormm_line_min *lm, int32_t move_number,
                                          void *stream) {
  int rc = check_objects(ag, ps, sc);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (sc == nullptr || lm == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "line advance needs a ScoreCard and a matching "
                                             "LineMinimization record");
  }
  stormm_ps_writer w;
  stormm_phasespace_writer(ps, &w);
  return line_advance(precision, ag, w, sc_view(sc), lm_view(lm), move_number, stream);
}

extern "C" int stormm_launch_line_advance_kit(int32_t precision, const stormm_synthesis *ag,
                                              const stormm_ps_writer *psw,
                                              const stormm_scorecard_writer *scw,
                                              const stormm_linmin_writer *lmw,
                                              int32_t move_number, void *stream) {
  if (psw == nullptr || scw == nullptr || lmw == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null writer");
  }
  ScView sc;
  LmView lm;
  int rc = sc_view(scw, &sc);
  if (rc == STORMM_OK) {
    rc = lm_view(lmw, &lm);
  }
  if (rc != STORMM_OK) {
    return rc;
  }
  return line_advance(precision, ag, *psw, sc, lm, move_number, stream);
}

// The reference's host loop, src/MolecularMechanics/hpc_minimization.cu:262-405, one fused
// evaluation per energy point.
static int minimize_stepwise(int32_t precision, const stormm_synthesis *ag,
                             const stormm_ps_writer &w, const stormm_minimize_controls *mc,
                             const ScView &sc, const LmView &lm, hipStream_t stream) {
  // The reference's per-cycle launch sequence (src/MolecularMechanics/hpc_minimization.cu:
  // 262-405).  The evaluations overwrite the force arrays and ScoreCard rows, which stands in
  // for the force and ScoreCard zeroing launches that precede them there.
  int step = 0;
  int rc = stormm_ps_writer_prime_conjugate_gradient(&w, ag->plan.padded_atoms, stream);
  if (rc != STORMM_OK) {
    return rc;
  }
  const int prec = (precision == STORMM_PREC_DOUBLE) ? 0 : 1;
  MMKit k = make_kit(ag, w, &sc);
  k.store_overwrite = 1;
  CgKit cg = make_cg_kit(ag, w, precision);
  cg.hist = sc.hist;
  const int terms = kTermNonbonded | kTermValence;
  const int cdcyc = mc->clash_damping_cycles;
  for (int cyc = 0; cyc < mc->total_cycles; cyc++) {
    // Clash damping of this cycle (hpc_minimization.cu:269-274) for all four evaluations.
    double cd = 0.0, cr = 0.0;
    if (cyc < cdcyc) {
      const double progress = static_cast<double>(cdcyc - cyc) / static_cast<double>(cdcyc);
      cd = mc->clash_minimum_distance * progress;
      cr = mc->clash_vdw_ratio * progress;
    }
    set_clash(k, precision, cd, cr);
    k.rst_step = minimizer_step(cyc, 0, cdcyc);
    HIP_TRY(launch_mm_eval(prec, k, terms, 1, 1, stream));
    step += 1;
    HIP_TRY(launch_conjugate_gradient(cg, step, mc->steepest_descent_cycles, stream));
    for (int move = 0; move < 4; move++) {
      if (move > 0) {
        k.rst_step = minimizer_step(cyc, move, cdcyc);
        HIP_TRY(launch_mm_eval(prec, k, terms, 0, 1, stream));
      }
      if ((rc = line_advance(precision, ag, w, sc, lm, move, stream)) != STORMM_OK) {
        return rc;
      }
    }
  }
  set_clash(k, precision, 0.0, 0.0);
  k.rst_step = minimizer_step(mc->total_cycles, 1, cdcyc);
  HIP_TRY(launch_mm_eval(prec, k, terms, 0, 1, stream));
  return STORMM_OK;
}

// The per-cycle launch sequence captured once in a HIP graph and replayed for every cycle.  The
// conjugate-gradient kernel reads each system's cycle count from device memory (so one captured
// cycle serves them all) and applies the RMS-force stop of the persistent minimizer; evaluations
// overwrite, so the graph holds kernels only.  With clash damping a second graph, whose
// evaluations take each system's damping schedule from the same cycle counts, serves the first
// clash_damping_cycles cycles.
static hipError_t capture_cycle(int prec, const MMKit &kf, const MMKit *ke, const CgKit &cg,
                                const LineKit &lk, int sd, hipStream_t cap,
                                hipGraph_t *graph, hipGraphExec_t *exec) {
  const int terms = kTermNonbonded | kTermValence;
  hipError_t err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) {
    return err;
  }
  hipError_t e2 = launch_mm_eval(prec, kf, terms, 1, 1, cap);
  if (e2 == hipSuccess) {
    e2 = launch_conjugate_gradient(cg, 0, sd, cap);
  }
  for (int move = 0; move < 4 && e2 == hipSuccess; move++) {
    if (move > 0) {
      e2 = launch_mm_eval(prec, ke[move - 1], terms, 0, 1, cap);
    }
    if (e2 == hipSuccess) {
      e2 = launch_line_advance(lk, move, cap);
    }
  }
  err = hipStreamEndCapture(cap, graph);
  if (err == hipSuccess) {
    err = e2;
  }
  if (err == hipSuccess) {
    err = hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0);
  }
  return err;
}

static int minimize_graph(int32_t precision, const stormm_synthesis *ag_c,
                          const stormm_ps_writer &w, const stormm_minimize_controls *mc,
                          const ScView &sc, const LmView &lm, hipStream_t stream) {
  stormm_synthesis *ag = const_cast<stormm_synthesis *>(ag_c);
  int rc = stormm_ps_writer_prime_conjugate_gradient(&w, ag->plan.padded_atoms, stream);
  if (rc != STORMM_OK) {
    return rc;
  }
  const int ns = ag->plan.nsys;
  HIP_TRY(hipMemsetAsync(ag->cycles_run.ptr, 0, ns * sizeof(int), stream));
  HIP_TRY(hipMemsetAsync(ag->stop_flags.ptr, 0, ns * sizeof(int), stream));
  HIP_TRY(hipMemsetAsync(ag->rms_force.ptr, 0, ns * sizeof(double), stream));
  const int prec = (precision == STORMM_PREC_DOUBLE) ? 0 : 1;
  MMKit k = make_kit(ag, w, &sc);
  k.store_overwrite = 1;
  set_clash(k, precision, 0.0, 0.0);
  CgKit cg = make_cg_kit(ag, w, precision);
  cg.hist = sc.hist;
  cg.steps = ag->cycles_run.as<int>();
  cg.done = ag->stop_flags.as<int>();
  cg.rms = ag->rms_force.as<double>();
  cg.conv_rms = mc->convergence_rms_force;
  cg.max_steps = mc->total_cycles;
  LineKit lk = make_line_kit(ag, w, sc, lm, precision);
  lk.done = ag->stop_flags.as<int>();
  const int terms = kTermNonbonded | kTermValence;
  const int sd = mc->steepest_descent_cycles;
  const int cdcyc = std::min(mc->clash_damping_cycles, mc->total_cycles);
  // Damped evaluations: the force evaluation reads the cycle count before the conjugate-gradient
  // kernel advances it, the three energy evaluations after.
  MMKit kf_clash = k, ke_clash = k;
  kf_clash.clash_variant = 1;
  kf_clash.clash_cycles = cdcyc;
  kf_clash.clash_step = ag->cycles_run.as<int>();
  kf_clash.clash_step_offset = 0;
  kf_clash.clash_r0 = mc->clash_minimum_distance;
  kf_clash.clash_ratio = mc->clash_vdw_ratio;
  ke_clash = kf_clash;
  ke_clash.clash_step_offset = -1;
  // Restraint steps from the same cycle counts: the force evaluation reads cycle i, the energy
  // evaluations i + 1.
  k.rst_step_sys = ag->cycles_run.as<int>();
  k.rst_clash_cycles = cdcyc;
  kf_clash.rst_step_sys = k.rst_step_sys;
  kf_clash.rst_clash_cycles = cdcyc;
  MMKit ke[3], ke_cl[3];
  for (int m = 0; m < 3; m++) {
    ke[m] = k;
    ke_cl[m] = ke_clash;
    ke[m].rst_cycle_offset = ke_cl[m].rst_cycle_offset = -1;
    ke[m].rst_move = ke_cl[m].rst_move = m + 1;
    ke_cl[m].rst_step_sys = k.rst_step_sys;
    ke_cl[m].rst_clash_cycles = cdcyc;
  }
  hipStream_t cap = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  hipGraph_t graph = nullptr, graph_clash = nullptr;
  hipGraphExec_t exec = nullptr, exec_clash = nullptr;
  hipError_t err = capture_cycle(prec, k, ke, cg, lk, sd, cap, &graph, &exec);
  if (err == hipSuccess && cdcyc > 0) {
    err = capture_cycle(prec, kf_clash, ke_cl, cg, lk, sd, cap, &graph_clash, &exec_clash);
  }
  for (int c = 0; c < mc->total_cycles && err == hipSuccess; c++) {
    err = hipGraphLaunch((c < cdcyc) ? exec_clash : exec, stream);
  }
  if (err == hipSuccess) {
    MMKit kfin = k;
    kfin.rst_step_sys = nullptr;
    kfin.rst_step = minimizer_step(mc->total_cycles, 1, cdcyc);
    err = launch_mm_eval(prec, kfin, terms, 0, 1, stream);
  }
  if (err == hipSuccess) {
    err = hipStreamSynchronize(stream);
  }
  for (hipGraphExec_t e : {exec, exec_clash}) {
    if (e != nullptr) {
      (void)hipGraphExecDestroy(e);
    }
  }
  for (hipGraph_t g : {graph, graph_clash}) {
    if (g != nullptr) {
      (void)hipGraphDestroy(g);
    }
  }
  (void)hipStreamDestroy(cap);
  HIP_TRY(err);
  return STORMM_OK;
}

// Per-system minimizer state of the fused cycle kernels: completed cycles, stop flags and the
// RMS force of the last direction update.
struct CycleState {
  int *steps = nullptr;
  int *done = nullptr;
  double *rms = nullptr;
};

// How the fused cycle kernels are issued.
enum class FusedIssue {
  kDirect,      // launched one by one on the caller's stream; the call returns at once
  kWholeGraph,  // the whole run captured in one HIP graph, launched once; returns when done
  kCycleGraph   // one cycle captured (plus a clash-damped one) and replayed; returns when done
};

// Launches of one fused cycle: the force evaluation (making the previous cycle's last line move
// first), the conjugate-gradient update, and three energy evaluations, each making the line move
// that follows the evaluation before it.
static hipError_t launch_fused_cycle(int prec, const MMKit &kf, const MMKit &ke,
                                     const CycleKit &ck, int sd_cycles, hipStream_t st) {
  hipError_t err = launch_min_cycle(prec, kf, ck, true, 3, st);
  if (err == hipSuccess) {
    err = launch_conjugate_gradient(ck.cg, 0, sd_cycles, st);
  }
  for (int move = 0; move < 3 && err == hipSuccess; move++) {
    err = launch_min_cycle(prec, ke, ck, false, move, st);
  }
  return err;
}

// Captures one fused cycle into an executable graph.
static hipError_t capture_fused_cycle(int prec, const MMKit &kf, const MMKit &ke,
                                      const CycleKit &ck, int sd_cycles, hipStream_t cap,
                                      hipGraph_t *graph, hipGraphExec_t *exec) {
  hipError_t err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) {
    return err;
  }
  const hipError_t e2 = launch_fused_cycle(prec, kf, ke, ck, sd_cycles, cap);
  err = hipStreamEndCapture(cap, graph);
  if (err == hipSuccess) {
    err = e2;
  }
  if (err == hipSuccess) {
    err = hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0);
  }
  return err;
}

// The reference's minimization loop (hpc_minimization.cu:262-405) with each line move made by
// the evaluation that follows it (launch_fused_cycle), then the final energy evaluation, which
// makes the last move.  Clash damping follows each system's cycle count in cs.steps, as in the
// graph replay, so one captured cycle serves every cycle of its kind.
static int minimize_fused(int32_t precision, const stormm_synthesis *ag,
                          const stormm_ps_writer &w, const stormm_minimize_controls *mc,
                          const ScView &sc, const LmView &lm, const CycleState &cs,
                          hipStream_t stream, FusedIssue how) {
  if (!ag->plan.rst.empty()) {
    return fail(STORMM_ERR_STATE, "the fused-cycle minimizer does not evaluate restraints; "
                                  "minimize a restrained synthesis with the per-step (0) or "
                                  "graph (2) mode");
  }
  if (w.gpos_bits > kMaxPlainGposBits || w.frc_bits > kMaxPlainFrcBits) {
    return fail(STORMM_ERR_STATE, "the fused-cycle minimizer takes at most 38 coordinate and 40 "
                                  "force bits; use the per-step (0) or graph (2) mode");
  }
  const int ns = ag->plan.nsys;
  const int prec = (precision == STORMM_PREC_DOUBLE) ? 0 : 1;
  MMKit k = make_kit(ag, w, &sc);
  k.store_overwrite = 1;
  set_clash(k, precision, 0.0, 0.0);
  CycleKit ck;
  ck.cg = make_cg_kit(ag, w, precision);
  ck.cg.hist = sc.hist;
  ck.cg.steps = cs.steps;
  ck.cg.done = cs.done;
  ck.cg.rms = cs.rms;
  ck.cg.conv_rms = mc->convergence_rms_force;
  ck.cg.max_steps = mc->total_cycles;
  ck.line = make_line_kit(ag, w, sc, lm, precision);
  ck.line.done = cs.done;
  const int sd = mc->steepest_descent_cycles;
  const int cdcyc = std::min(mc->clash_damping_cycles, mc->total_cycles);
  // Damped evaluations read the cycle count before this cycle's direction update advances it
  // (force evaluation) or after (energy evaluations).
  MMKit kf_clash = k;
  kf_clash.clash_variant = 1;
  kf_clash.clash_cycles = cdcyc;
  kf_clash.clash_step = cs.steps;
  kf_clash.clash_step_offset = 0;
  kf_clash.clash_r0 = mc->clash_minimum_distance;
  kf_clash.clash_ratio = mc->clash_vdw_ratio;
  MMKit ke_clash = kf_clash;
  ke_clash.clash_step_offset = -1;
  MMKit kfin = k;
  kfin.rst_step = minimizer_step(mc->total_cycles, 1, cdcyc);
  const int terms = kTermNonbonded | kTermValence;
  hipStream_t cap = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipStream_t issue = stream;
  hipError_t err = hipSuccess;
  if (how != FusedIssue::kDirect) {
    HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  }
  if (how == FusedIssue::kWholeGraph) {
    err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    issue = cap;
  }
  if (err == hipSuccess) {
    err = (stormm_ps_writer_prime_conjugate_gradient(&w, ag->plan.padded_atoms, issue) ==
           STORMM_OK) ? hipSuccess : hipErrorInvalidValue;
  }
  if (err == hipSuccess) {
    err = launch_fill_int(cs.steps, cs.done, cs.done, static_cast<size_t>(ns), issue);
  }
  if (err == hipSuccess) {
    err = launch_fill(reinterpret_cast<long long *>(cs.rms), nullptr, nullptr,
                      static_cast<size_t>(ns), issue);
  }
  if (how == FusedIssue::kCycleGraph) {
    if (err == hipSuccess) {
      err = capture_fused_cycle(prec, k, k, ck, sd, cap, &graph[0], &exec[0]);
    }
    if (err == hipSuccess && cdcyc > 0) {
      err = capture_fused_cycle(prec, kf_clash, ke_clash, ck, sd, cap, &graph[1], &exec[1]);
    }
    for (int c = 0; c < mc->total_cycles && err == hipSuccess; c++) {
      err = hipGraphLaunch(exec[(c < cdcyc) ? 1 : 0], stream);
    }
  }
  else {
    for (int c = 0; c < mc->total_cycles && err == hipSuccess; c++) {
      const bool damped = (c < cdcyc);
      err = launch_fused_cycle(prec, damped ? kf_clash : k, damped ? ke_clash : k, ck, sd, issue);
    }
  }
  if (err == hipSuccess) {
    // The final energies, after the last cycle's last move.
    err = (mc->total_cycles > 0) ? launch_min_cycle(prec, kfin, ck, false, 3, issue)
                                 : launch_mm_eval(prec, kfin, terms, 0, 1, issue);
  }
  if (how == FusedIssue::kWholeGraph) {
    const hipError_t e1 = hipStreamEndCapture(cap, &graph[0]);
    if (err == hipSuccess) {
      err = e1;
    }
    if (err == hipSuccess) {
      err = hipGraphInstantiate(&exec[0], graph[0], nullptr, nullptr, 0);
    }
    if (err == hipSuccess) {
      err = hipGraphLaunch(exec[0], stream);
    }
  }
  if (how != FusedIssue::kDirect) {
    if (err == hipSuccess) {
      err = hipStreamSynchronize(stream);
    }
    for (int g = 0; g < 2; g++) {
      if (exec[g] != nullptr) {
        (void)hipGraphExecDestroy(exec[g]);
      }
      if (graph[g] != nullptr) {
        (void)hipGraphDestroy(graph[g]);
      }
    }
    (void)hipStreamDestroy(cap);
  }
  HIP_TRY(err);
  return STORMM_OK;
}

static int minimize(int32_t precision, const stormm_synthesis *ag, const stormm_ps_writer &w,
                    const stormm_minimize_controls *mincon, const ScView &sc, const LmView &lm,
                    int32_t persistent, void *stream) {
  int rc = check_writer(ag, &w, precision, &sc, true);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (mincon == nullptr || sc.acc == nullptr || lm.nsys != ag->plan.nsys) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "minimization needs controls, a ScoreCard and a "
                                             "matching LineMinimization record");
  }
  if ((rc = check_mincon(mincon)) != STORMM_OK) {
    return rc;
  }
  if (persistent < 0 || persistent > 4) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "minimization mode must be 0, 1, 2, 3 or 4");
  }
  ScView sch = sc;
  if ((rc = prepare_history(mincon, &sch)) != STORMM_OK) {
    return rc;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  stormm_synthesis *agm = const_cast<stormm_synthesis *>(ag);
  if (persistent == 0) {
    rc = minimize_stepwise(precision, ag, w, mincon, sch, lm, st);
  }
  else if (persistent == 2) {
    rc = minimize_graph(precision, ag, w, mincon, sch, lm, st);
  }
  else {
    CycleState cs;
    cs.steps = agm->cycles_run.as<int>();
    cs.done = agm->stop_flags.as<int>();
    cs.rms = agm->rms_force.as<double>();
    const FusedIssue how = (persistent == 1) ? FusedIssue::kWholeGraph :
                           (persistent == 3) ? FusedIssue::kCycleGraph : FusedIssue::kDirect;
    rc = minimize_fused(precision, ag, w, mincon, sch, lm, cs, st, how);
  }
  if (rc == STORMM_OK) {
    rc = finish_history(mincon, sch, (persistent == 0) ? nullptr : agm->cycles_run.as<int>(),
                        st);
  }
  if (rc == STORMM_OK && persistent == 4) {
    HIP_TRY(hipStreamSynchronize(st));
  }
  return rc;
}

extern "C" int stormm_launch_minimization(int32_t precision, const stormm_synthesis *ag,
                                          stormm_phasespace *ps,
                                          const stormm_minimize_controls *mincon,
                                          stormm_scorecard *sc, stormm_line_min *lm,
                                          int32_t persistent, void *stream) {
  int rc = check_objects(ag, ps, sc);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (mincon == nullptr || sc == nullptr || lm == nullptr || lm->nsys != ag->plan.nsys) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "minimization needs controls, a ScoreCard and a "
                                             "matching LineMinimization record");
  }
  stormm_ps_writer w;
  stormm_phasespace_writer(ps, &w);
  return minimize(precision, ag, w, mincon, sc_view(sc), lm_view(lm), persistent, stream);
}

extern "C" int stormm_launch_minimization_kit(int32_t precision, const stormm_synthesis *ag,
                                              const stormm_ps_writer *psw,
                                              const stormm_minimize_controls *mincon,
                                              const stormm_scorecard_writer *scw,
                                              const stormm_linmin_writer *lmw,
                                              int32_t persistent, void *stream) {
  if (psw == nullptr || scw == nullptr || lmw == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null writer");
  }
  ScView sc;
  LmView lm;
  int rc = sc_view(scw, &sc);
  if (rc == STORMM_OK) {
    rc = lm_view(lmw, &lm);
  }
  if (rc != STORMM_OK) {
    return rc;
  }
  return minimize(precision, ag, *psw, mincon, sc, lm, persistent, stream);
}

// The fused-cycle minimizer on caller-owned arrays with its per-system state in the library's
// static device array: no device allocation, no graph, no stream creation and no
// synchronization; the launches are queued on `stream` and the call returns.
extern "C" int stormm_launch_minimization_kit_noalloc(int32_t precision,
                                                      const stormm_synthesis *ag,
                                                      const stormm_ps_writer *psw,
                                                      const stormm_minimize_controls *mincon,
                                                      const stormm_scorecard_writer *scw,
                                                      const stormm_linmin_writer *lmw,
                                                      void *stream) {
  if (psw == nullptr || scw == nullptr || lmw == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null writer");
  }
  ScView sc;
  LmView lm;
  int rc = sc_view(scw, &sc);
  if (rc == STORMM_OK) {
    rc = lm_view(lmw, &lm);
  }
  if (rc == STORMM_OK) {
    rc = check_writer(ag, psw, precision, &sc, true);
  }
  if (rc != STORMM_OK) {
    return rc;
  }
  if (mincon == nullptr || sc.acc == nullptr || lm.nsys != ag->plan.nsys) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "minimization needs controls, a ScoreCard and a "
                                             "matching LineMinimization record");
  }
  if ((rc = check_mincon(mincon)) != STORMM_OK) {
    return rc;
  }
  if (ag->plan.nsys > kNoAllocMaxSystems) {
    return fail(STORMM_ERR_CAPACITY, "the allocation-free minimization takes at most " +
                                         std::to_string(kNoAllocMaxSystems) + " systems (" +
                                         std::to_string(ag->plan.nsys) + " given)");
  }
  // A writer's history is the caller's: checked for room, never grown.
  if ((rc = prepare_history(mincon, &sc)) != STORMM_OK) {
    return rc;
  }
  void *base = nullptr;
  HIP_TRY(noalloc_scratch(&base));
  CycleState cs;
  cs.steps = static_cast<int *>(base);
  cs.done = cs.steps + kNoAllocMaxSystems;
  cs.rms = reinterpret_cast<double *>(static_cast<unsigned char *>(base) +
                                      (static_cast<size_t>(kNoAllocMaxSystems) * 8));
  hipStream_t st = static_cast<hipStream_t>(stream);
  rc = minimize_fused(precision, ag, *psw, mincon, sc, lm, cs, st, FusedIssue::kDirect);
  if (rc == STORMM_OK) {
    rc = finish_history(mincon, sc, cs.steps, st);
  }
  return rc;
}

extern "C" int stormm_check_ps_writer(const stormm_synthesis *ag, const stormm_ps_writer *psw) {
  if (ag == nullptr || psw == nullptr || psw->atom_starts == nullptr ||
      psw->atom_counts == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null synthesis, writer or layout arrays");
  }
  const SynthesisPlan &P = ag->plan;
  if (psw->system_count != P.nsys) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "the writer holds " +
                                                 std::to_string(psw->system_count) +
                                                 " systems, the synthesis " +
                                                 std::to_string(P.nsys));
  }
  std::vector<int> starts(P.nsys), counts(P.nsys);
  HIP_TRY(hipMemcpy(starts.data(), psw->atom_starts, P.nsys * sizeof(int),
                    hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(counts.data(), psw->atom_counts, P.nsys * sizeof(int),
                    hipMemcpyDeviceToHost));
  for (int s = 0; s < P.nsys; s++) {
    if (starts[s] != P.atom_starts[s] || counts[s] != P.atom_counts[s]) {
      return fail(STORMM_ERR_INVALID_ARGUMENT,
                  "system " + std::to_string(s) + " lies at atoms " + std::to_string(starts[s]) +
                      " + " + std::to_string(counts[s]) + " in the writer, at " +
                      std::to_string(P.atom_starts[s]) + " + " +
                      std::to_string(P.atom_counts[s]) + " in the synthesis");
    }
  }
  return STORMM_OK;
}

extern "C" int stormm_minimization_status(const stormm_synthesis *ag, int32_t *cycles_run,
                                          double *rms_force) {
  if (ag == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null synthesis");
  }
  if (cycles_run) {
    HIP_TRY(hipMemcpy(cycles_run, ag->cycles_run.ptr, ag->plan.nsys * sizeof(int),
                      hipMemcpyDeviceToHost));
  }
  if (rms_force) {
    HIP_TRY(hipMemcpy(rms_force, ag->rms_force.ptr, ag->plan.nsys * sizeof(double),
                      hipMemcpyDeviceToHost));
  }
  return STORMM_OK;
}

extern "C" int stormm_synthesis_status(const stormm_synthesis *ag, int32_t *status,
                                       int32_t clear) {
  if (ag == nullptr || status == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  HIP_TRY(hipMemcpy(status, ag->status.ptr, sizeof(int), hipMemcpyDeviceToHost));
  if (clear) {
    HIP_TRY(hipMemset(ag->status.ptr, 0, sizeof(int)));
  }
  return STORMM_OK;
}

extern "C" int stormm_time_mm_evaluation(int32_t precision, const stormm_synthesis *ag,
                                         stormm_phasespace *ps, stormm_scorecard *sc,
                                         int32_t eval_force, int32_t repeats, void *stream,
                                         double *mean_ms) {
  return stormm_time_launch(precision, ag, ps, sc, STORMM_TERMS_ALL, eval_force, repeats,
                            stream, mean_ms);
}

extern "C" int stormm_time_launch(int32_t precision, const stormm_synthesis *ag,
                                  stormm_phasespace *ps, stormm_scorecard *sc, int32_t terms,
                                  int32_t eval_force, int32_t repeats, void *stream,
                                  double *mean_ms) {
  if (terms < STORMM_TERMS_NONBONDED || terms > STORMM_TERMS_ALL) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "terms must be STORMM_TERMS_NONBONDED, "
                                             "STORMM_TERMS_VALENCE or both");
  }
  int rc = check_objects(ag, ps, sc);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (repeats <= 0 || mean_ms == nullptr || sc == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid timing request");
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  stormm_ps_writer w;
  stormm_phasespace_writer(ps, &w);
  const ScView scv = sc_view(sc);
  if ((rc = check_writer(ag, &w, precision, &scv, false)) != STORMM_OK) {
    return rc;
  }
  const MMKit k = make_kit(ag, w, &scv);
  const int prec = (precision == STORMM_PREC_DOUBLE) ? 0 : 1;
  HIP_TRY(hipEventRecord(e0, st));
  for (int r = 0; r < repeats; r++) {
    HIP_TRY(launch_mm_eval(prec, k, terms, eval_force, 1, st));
  }
  HIP_TRY(hipEventRecord(e1, st));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  HIP_TRY(hipEventDestroy(e0));
  HIP_TRY(hipEventDestroy(e1));
  *mean_ms = static_cast<double>(ms) / repeats;
  return STORMM_OK;
}

extern "C" int stormm_autotune_launch_shape(const stormm_topology_batch *batch,
                                            int32_t precision, const double *coordinates,
                                            stormm_launch_shape *best, double *best_ms) {
  if (batch == nullptr || best == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  int maxn = 0;
  for (int s = 0; s < batch->system_count; s++) {
    maxn = std::max(maxn, batch->atom_counts[s]);
  }
  const int caps[4] = {128, 256, 384, 512};
  const int per_cu[4] = {1, 2, 3, 4};
  const int blocks[3] = {256, 512, 1024};
  double best_t = 1.0e30;
  stormm_launch_shape best_shape = {512, 2, kEvalBlock};
  // Timing coordinates: the caller's (the near / far mix of the Born-radius pairs depends on
  // the geometry), or else a loose cubic lattice per system (1.6 A spacing).
  std::vector<double> xyz;
  if (coordinates == nullptr) {
    for (int s = 0; s < batch->system_count; s++) {
      for (int i = 0; i < batch->atom_counts[s]; i++) {
        xyz.push_back(1.6 * (i % 8));
        xyz.push_back(1.6 * ((i / 8) % 8));
        xyz.push_back(1.6 * (i / 64));
      }
    }
  }
  for (int cap : caps) {
    if (cap < round_up(maxn, kTile)) {
      continue;
    }
    for (int pc : per_cu) {
     for (int bt : blocks) {
      stormm_launch_shape shape = {cap, pc, bt};
      stormm_synthesis *ag = nullptr;
      int rc = stormm_synthesis_create(batch, &shape, &ag);
      if (rc != STORMM_OK) {
        return rc;
      }
      stormm_phasespace *ps = nullptr;
      stormm_scorecard *sc = nullptr;
      rc = stormm_phasespace_create(ag, 32, 23, &ps);
      if (rc == STORMM_OK) {
        rc = stormm_scorecard_create(ag->plan.nsys, 26, &sc);
      }
      if (rc == STORMM_OK) {
        rc = stormm_phasespace_upload_coordinates(ps, coordinates ? coordinates : xyz.data());
      }
      double ms = 0.0;
      if (rc == STORMM_OK) {
        rc = stormm_time_mm_evaluation(precision, ag, ps, sc, 1, 2, nullptr, &ms);
      }
      if (rc == STORMM_OK) {
        rc = stormm_time_mm_evaluation(precision, ag, ps, sc, 1, 5, nullptr, &ms);
      }
      stormm_scorecard_destroy(sc);
      stormm_phasespace_destroy(ps);
      stormm_synthesis_destroy(ag);
      if (rc != STORMM_OK) {
        return rc;
      }
      if (ms < best_t) {
        best_t = ms;
        best_shape = shape;
      }
     }
    }
  }
  *best = best_shape;
  if (best_ms) {
    *best_ms = best_t;
  }
  return STORMM_OK;
}
