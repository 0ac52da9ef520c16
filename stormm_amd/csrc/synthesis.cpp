// Host side of the C-ABI: AtomGraphSynthesis construction (parameter layout, Lennard-Jones mode
// detection, work-unit packing, tile tasks, exclusion masks and valence term streams), the
// PhaseSpaceSynthesis / ScoreCard / controls objects, and the launch entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/stormm_amd.h"
#include "host_common.h"
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
};

struct stormm_scorecard {
  int nsys = 0, nrg_bits = 26, samples = 0;
  DevMem inst, running, squared;
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
extern "C" int stormm_phasespace_create(const stormm_synthesis *ag, int32_t gpos_bits,
                                        int32_t frc_bits, stormm_phasespace **out) {
  if (ag == nullptr || out == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  if (gpos_bits < 24 || gpos_bits > 38) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "gpos_bits must lie in [24, 38] (64-bit "
                                             "coordinates without overflow words)");
  }
  if (frc_bits < 18 || frc_bits > 40) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "frc_bits must lie in [18, 40]");
  }
  std::unique_ptr<stormm_phasespace> ps(new stormm_phasespace());
  ps->ag = ag;
  ps->gpos_bits = gpos_bits;
  ps->frc_bits = frc_bits;
  const size_t n = static_cast<size_t>(ag->plan.padded_atoms) * sizeof(long long);
  DevMem *arrs[12] = {&ps->xcrd, &ps->ycrd, &ps->zcrd, &ps->xfrc, &ps->yfrc, &ps->zfrc,
                      &ps->xprv, &ps->yprv, &ps->zprv, &ps->xcg, &ps->ycg, &ps->zcg};
  for (DevMem *d : arrs) {
    HIP_TRY(d->alloc(n));
    HIP_TRY(hipMemset(d->ptr, 0, n));
  }
  *out = ps.release();
  return STORMM_OK;
}

extern "C" int stormm_phasespace_destroy(stormm_phasespace *ps) {
  delete ps;
  return STORMM_OK;
}

static int total_atoms(const SynthesisPlan &P) {
  int t = 0;
  for (int s = 0; s < P.nsys; s++) {
    t += P.atom_counts[s];
  }
  return t;
}

extern "C" int stormm_phasespace_upload_fixed(stormm_phasespace *ps, const int64_t *xyz) {
  if (ps == nullptr || xyz == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  const SynthesisPlan &P = ps->ag->plan;
  std::vector<long long> hx(P.padded_atoms, 0), hy(P.padded_atoms, 0), hz(P.padded_atoms, 0);
  int c = 0;
  for (int s = 0; s < P.nsys; s++) {
    for (int i = 0; i < P.atom_counts[s]; i++) {
      const int g = P.atom_starts[s] + i;
      hx[g] = xyz[3 * c];
      hy[g] = xyz[3 * c + 1];
      hz[g] = xyz[3 * c + 2];
      c++;
    }
  }
  const size_t n = hx.size() * sizeof(long long);
  HIP_TRY(hipMemcpy(ps->xcrd.ptr, hx.data(), n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ps->ycrd.ptr, hy.data(), n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ps->zcrd.ptr, hz.data(), n, hipMemcpyHostToDevice));
  return STORMM_OK;
}

extern "C" int stormm_phasespace_upload_coordinates(stormm_phasespace *ps, const double *xyz) {
  if (ps == nullptr || xyz == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  const int tot = total_atoms(ps->ag->plan);
  std::vector<int64_t> fx(static_cast<size_t>(3) * tot);
  const double scale = std::ldexp(1.0, ps->gpos_bits);
  for (size_t k = 0; k < fx.size(); k++) {
    fx[k] = std::llround(xyz[k] * scale);
  }
  return stormm_phasespace_upload_fixed(ps, fx.data());
}

extern "C" int stormm_phasespace_download_fixed(const stormm_phasespace *ps, int64_t *xyz,
                                                int64_t *fxyz) {
  if (ps == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  const SynthesisPlan &P = ps->ag->plan;
  const size_t n = static_cast<size_t>(P.padded_atoms) * sizeof(long long);
  std::vector<long long> a(P.padded_atoms), b(P.padded_atoms), c(P.padded_atoms);
  for (int pass = 0; pass < 2; pass++) {
    int64_t *dst = (pass == 0) ? xyz : fxyz;
    if (dst == nullptr) {
      continue;
    }
    const DevMem &dx = (pass == 0) ? ps->xcrd : ps->xfrc;
    const DevMem &dy = (pass == 0) ? ps->ycrd : ps->yfrc;
    const DevMem &dz = (pass == 0) ? ps->zcrd : ps->zfrc;
    HIP_TRY(hipMemcpy(a.data(), dx.ptr, n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(b.data(), dy.ptr, n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(c.data(), dz.ptr, n, hipMemcpyDeviceToHost));
    int k = 0;
    for (int s = 0; s < P.nsys; s++) {
      for (int i = 0; i < P.atom_counts[s]; i++) {
        const int g = P.atom_starts[s] + i;
        dst[3 * k] = a[g];
        dst[3 * k + 1] = b[g];
        dst[3 * k + 2] = c[g];
        k++;
      }
    }
  }
  return STORMM_OK;
}

extern "C" int stormm_phasespace_download_coordinates(const stormm_phasespace *ps, double *xyz) {
  if (ps == nullptr || xyz == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  const int tot = total_atoms(ps->ag->plan);
  std::vector<int64_t> fx(static_cast<size_t>(3) * tot);
  int rc = stormm_phasespace_download_fixed(ps, fx.data(), nullptr);
  if (rc != STORMM_OK) {
    return rc;
  }
  const double inv = std::ldexp(1.0, -ps->gpos_bits);
  for (size_t k = 0; k < fx.size(); k++) {
    xyz[k] = static_cast<double>(fx[k]) * inv;
  }
  return STORMM_OK;
}

extern "C" int stormm_phasespace_download_forces(const stormm_phasespace *ps, double *fxyz) {
  if (ps == nullptr || fxyz == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  const int tot = total_atoms(ps->ag->plan);
  std::vector<int64_t> ff(static_cast<size_t>(3) * tot);
  int rc = stormm_phasespace_download_fixed(ps, nullptr, ff.data());
  if (rc != STORMM_OK) {
    return rc;
  }
  const double inv = std::ldexp(1.0, -ps->frc_bits);
  for (size_t k = 0; k < ff.size(); k++) {
    fxyz[k] = static_cast<double>(ff[k]) * inv;
  }
  return STORMM_OK;
}

// Systems [lo, hi) occupy one contiguous span of device atoms, system after system with each
// system padded to the wave width; the helpers below move that span in one copy per axis.
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

extern "C" int stormm_phasespace_upload_systems(stormm_phasespace *ps, int32_t lo, int32_t hi,
                                                const double *xyz) {
  int b = 0, e = 0;
  int rc = system_span(ps, lo, hi, &b, &e);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (xyz == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null coordinates");
  }
  const SynthesisPlan &P = ps->ag->plan;
  const double scale = std::ldexp(1.0, ps->gpos_bits);
  std::vector<long long> hx(e - b, 0), hy(e - b, 0), hz(e - b, 0);
  size_t c = 0;
  for (int s = lo; s < hi; s++) {
    for (int i = 0; i < P.atom_counts[s]; i++) {
      const int g = P.atom_starts[s] + i - b;
      hx[g] = std::llround(xyz[3 * c] * scale);
      hy[g] = std::llround(xyz[3 * c + 1] * scale);
      hz[g] = std::llround(xyz[3 * c + 2] * scale);
      c++;
    }
  }
  const size_t n = hx.size() * sizeof(long long);
  HIP_TRY(hipMemcpy(ps->xcrd.as<long long>() + b, hx.data(), n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ps->ycrd.as<long long>() + b, hy.data(), n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ps->zcrd.as<long long>() + b, hz.data(), n, hipMemcpyHostToDevice));
  return STORMM_OK;
}

extern "C" int stormm_phasespace_download_systems(const stormm_phasespace *ps, int32_t lo,
                                                  int32_t hi, double *xyz, double *fxyz) {
  int b = 0, e = 0;
  int rc = system_span(ps, lo, hi, &b, &e);
  if (rc != STORMM_OK) {
    return rc;
  }
  const SynthesisPlan &P = ps->ag->plan;
  const size_t n = static_cast<size_t>(e - b) * sizeof(long long);
  std::vector<long long> a(e - b), bb(e - b), cc(e - b);
  for (int pass = 0; pass < 2; pass++) {
    double *dst = (pass == 0) ? xyz : fxyz;
    if (dst == nullptr) {
      continue;
    }
    const DevMem &dx = (pass == 0) ? ps->xcrd : ps->xfrc;
    const DevMem &dy = (pass == 0) ? ps->ycrd : ps->yfrc;
    const DevMem &dz = (pass == 0) ? ps->zcrd : ps->zfrc;
    HIP_TRY(hipMemcpy(a.data(), dx.as<long long>() + b, n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(bb.data(), dy.as<long long>() + b, n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cc.data(), dz.as<long long>() + b, n, hipMemcpyDeviceToHost));
    const double inv = std::ldexp(1.0, (pass == 0) ? -ps->gpos_bits : -ps->frc_bits);
    size_t k = 0;
    for (int s = lo; s < hi; s++) {
      for (int i = 0; i < P.atom_counts[s]; i++) {
        const int g = P.atom_starts[s] + i - b;
        dst[3 * k] = static_cast<double>(a[g]) * inv;
        dst[3 * k + 1] = static_cast<double>(bb[g]) * inv;
        dst[3 * k + 2] = static_cast<double>(cc[g]) * inv;
        k++;
      }
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
  return STORMM_OK;
}

extern "C" int stormm_phasespace_initialize_forces(stormm_phasespace *ps, void *stream) {
  if (ps == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  HIP_TRY(launch_fill(ps->xfrc.as<long long>(), ps->yfrc.as<long long>(),
                      ps->zfrc.as<long long>(), ps->ag->plan.padded_atoms,
                      static_cast<hipStream_t>(stream)));
  return STORMM_OK;
}

extern "C" int stormm_phasespace_prime_conjugate_gradient(stormm_phasespace *ps, void *stream) {
  if (ps == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  const size_t n = ps->ag->plan.padded_atoms;
  HIP_TRY(launch_fill(ps->xprv.as<long long>(), ps->yprv.as<long long>(),
                      ps->zprv.as<long long>(), n, static_cast<hipStream_t>(stream)));
  HIP_TRY(launch_fill(ps->xcg.as<long long>(), ps->ycg.as<long long>(), ps->zcg.as<long long>(),
                      n, static_cast<hipStream_t>(stream)));
  return STORMM_OK;
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

extern "C" int stormm_scorecard_commit(stormm_scorecard *sc, void *stream) {
  if (sc == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  HIP_TRY(launch_scorecard_commit(sc->inst.as<long long>(), sc->running.as<double>(),
                                  sc->squared.as<double>(), sc->nsys,
                                  std::ldexp(1.0, -sc->nrg_bits),
                                  static_cast<hipStream_t>(stream)));
  sc->samples += 1;
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

extern "C" int stormm_line_min_download(const stormm_line_min *lm, double *values9xN) {
  if (lm == nullptr || values9xN == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  HIP_TRY(hipMemcpy(values9xN, lm->vals.ptr, static_cast<size_t>(9) * lm->nsys * sizeof(double),
                    hipMemcpyDeviceToHost));
  return STORMM_OK;
}

// ---- Launches ------------------------------------------------------------------------------
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

static MMKit make_kit(const stormm_synthesis *ag, stormm_phasespace *ps, stormm_scorecard *sc) {
  MMKit k = ag->kit;
  k.gpos_bits = ps->gpos_bits;
  k.frc_bits = ps->frc_bits;
  k.nrg_bits = (sc != nullptr) ? sc->nrg_bits : 26;
  k.xcrd = ps->xcrd.as<long long>();
  k.ycrd = ps->ycrd.as<long long>();
  k.zcrd = ps->zcrd.as<long long>();
  k.xfrc = ps->xfrc.as<long long>();
  k.yfrc = ps->yfrc.as<long long>();
  k.zfrc = ps->zfrc.as<long long>();
  k.scorecard = (sc != nullptr) ? sc->inst.as<long long>() : nullptr;
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
  return STORMM_OK;
}

static int launch_terms(int32_t precision, const stormm_synthesis *ag, stormm_phasespace *ps,
                        stormm_scorecard *sc, int terms, int32_t eval_force, int32_t eval_energy,
                        double clash_d, double clash_r, int step, void *stream) {
  int rc = check_objects(ag, ps, sc);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (precision != STORMM_PREC_SINGLE && precision != STORMM_PREC_DOUBLE) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "precision must be STORMM_PREC_SINGLE or "
                                             "STORMM_PREC_DOUBLE");
  }
  if (sc == nullptr && eval_energy) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "energy evaluation needs a ScoreCard");
  }
  if (!(clash_d >= 0.0) || !(clash_r >= 0.0)) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "clash parameters must be non-negative");
  }
  MMKit k = make_kit(ag, ps, sc);
  set_clash(k, precision, clash_d, clash_r);
  k.rst_step = step;
  HIP_TRY(launch_mm_eval(precision == STORMM_PREC_DOUBLE ? 0 : 1, k, terms, eval_force ? 1 : 0,
                         eval_energy ? 1 : 0, static_cast<hipStream_t>(stream)));
  return STORMM_OK;
}

extern "C" int stormm_launch_nonbonded(int32_t precision, const stormm_synthesis *ag,
                                       stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                       stormm_scorecard *sc, int32_t eval_force,
                                       int32_t eval_energy, double clash_minimum_distance,
                                       double clash_ratio, void *stream) {
  return launch_terms(precision, ag, ps, sc, kTermNonbonded, eval_force, eval_energy,
                      clash_minimum_distance, clash_ratio,
                      (ctrl != nullptr) ? ctrl->step : 0, stream);
}

extern "C" int stormm_launch_valence(int32_t precision, const stormm_synthesis *ag,
                                     stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                     stormm_scorecard *sc, int32_t eval_force,
                                     int32_t eval_energy, double clash_minimum_distance,
                                     double clash_ratio, void *stream) {
  return launch_terms(precision, ag, ps, sc, kTermValence, eval_force, eval_energy,
                      clash_minimum_distance, clash_ratio,
                      (ctrl != nullptr) ? ctrl->step : 0, stream);
}

extern "C" int stormm_launch_mm_evaluation(int32_t precision, const stormm_synthesis *ag,
                                           stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                           stormm_scorecard *sc, int32_t eval_force,
                                           int32_t eval_energy, double clash_minimum_distance,
                                           double clash_ratio, void *stream) {
  return launch_terms(precision, ag, ps, sc, kTermNonbonded | kTermValence, eval_force,
                      eval_energy, clash_minimum_distance, clash_ratio,
                      (ctrl != nullptr) ? ctrl->step : 0, stream);
}

static CgKit make_cg_kit(const stormm_synthesis *ag, stormm_phasespace *ps, int precision) {
  CgKit k;
  std::memset(&k, 0, sizeof(k));
  k.nsys = ag->plan.nsys;
  k.sys_start = ag->atom_starts.as<int>();
  k.sys_natom = ag->atom_counts.as<int>();
  k.xfrc = ps->xfrc.as<long long>();
  k.yfrc = ps->yfrc.as<long long>();
  k.zfrc = ps->zfrc.as<long long>();
  k.xprv = ps->xprv.as<long long>();
  k.yprv = ps->yprv.as<long long>();
  k.zprv = ps->zprv.as<long long>();
  k.xcg = ps->xcg.as<long long>();
  k.ycg = ps->ycg.as<long long>();
  k.zcg = ps->zcg.as<long long>();
  k.inv_frc_scale = std::ldexp(1.0, -ps->frc_bits);
  k.dp_mode = (precision == STORMM_PREC_DOUBLE) ? 1 : 0;
  return k;
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
  HIP_TRY(launch_conjugate_gradient(make_cg_kit(ag, ps, precision), ctrl->step,
                                    ctrl->mc.steepest_descent_cycles,
                                    static_cast<hipStream_t>(stream)));
  return STORMM_OK;
}

static LineKit make_line_kit(const stormm_synthesis *ag, stormm_phasespace *ps,
                             stormm_scorecard *sc, stormm_line_min *lm, int precision) {
  LineKit k;
  std::memset(&k, 0, sizeof(k));
  k.nsys = ag->plan.nsys;
  k.sys_start = ag->atom_starts.as<int>();
  k.sys_natom = ag->atom_counts.as<int>();
  k.xcrd = ps->xcrd.as<long long>();
  k.ycrd = ps->ycrd.as<long long>();
  k.zcrd = ps->zcrd.as<long long>();
  k.xfrc = ps->xfrc.as<long long>();
  k.yfrc = ps->yfrc.as<long long>();
  k.zfrc = ps->zfrc.as<long long>();
  k.scorecard = sc->inst.as<long long>();
  k.inv_nrg_scale = std::ldexp(1.0, -sc->nrg_bits);
  k.inv_frc_scale = std::ldexp(1.0, -ps->frc_bits);
  k.gpos_scale = std::ldexp(1.0, ps->gpos_bits);
  double *v = lm->vals.as<double>();
  const int n = lm->nsys;
  k.l_move = v;
  k.s_move = v + n;
  k.mfac_a = v + 2 * n;
  k.mfac_b = v + 3 * n;
  k.mfac_c = v + 4 * n;
  k.nrg_a = v + 5 * n;
  k.nrg_b = v + 6 * n;
  k.nrg_c = v + 7 * n;
  k.nrg_d = v + 8 * n;
  k.dp_mode = (precision == STORMM_PREC_DOUBLE) ? 1 : 0;
  return k;
}

extern "C" int stormm_launch_line_advance(int32_t precision, const stormm_synthesis *ag,
                                          stormm_phasespace *ps, stormm_scorecard *sc,
                                          stormm_line_min *lm, int32_t move_number,
                                          void *stream) {
  int rc = check_objects(ag, ps, sc);
  if (rc != STORMM_OK) {
    return rc;
  }
  if (sc == nullptr || lm == nullptr || lm->nsys != ag->plan.nsys) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "line advance needs a ScoreCard and a matching "
                                             "LineMinimization record");
  }
  if (move_number < 0 || move_number > 3) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "move_number must be 0..3");
  }
  HIP_TRY(launch_line_advance(make_line_kit(ag, ps, sc, lm, precision), move_number,
                              static_cast<hipStream_t>(stream)));
  return STORMM_OK;
}

// The reference's host loop, src/MolecularMechanics/hpc_minimization.cu:262-405, one fused
// evaluation per energy point.
static int minimize_stepwise(int32_t precision, const stormm_synthesis *ag, stormm_phasespace *ps,
                             const stormm_minimize_controls *mc, stormm_scorecard *sc,
                             stormm_line_min *lm, hipStream_t stream) {
  // The reference's per-cycle launch sequence (src/MolecularMechanics/hpc_minimization.cu:
  // 262-405).  The evaluations overwrite the force arrays and ScoreCard rows, which stands in
  // for the force and ScoreCard zeroing launches that precede them there.
  stormm_mm_controls ctrl;
  ctrl.mc = *mc;
  ctrl.step = 0;
  int rc = stormm_phasespace_prime_conjugate_gradient(ps, stream);
  if (rc != STORMM_OK) {
    return rc;
  }
  const int prec = (precision == STORMM_PREC_DOUBLE) ? 0 : 1;
  MMKit k = make_kit(ag, ps, sc);
  k.store_overwrite = 1;
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
    ctrl.step += 1;
    if ((rc = stormm_launch_conjugate_gradient(precision, ag, ps, &ctrl, stream)) != STORMM_OK) {
      return rc;
    }
    for (int move = 0; move < 4; move++) {
      if (move > 0) {
        k.rst_step = minimizer_step(cyc, move, cdcyc);
        HIP_TRY(launch_mm_eval(prec, k, terms, 0, 1, stream));
      }
      if ((rc = stormm_launch_line_advance(precision, ag, ps, sc, lm, move, stream)) !=
          STORMM_OK) {
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

static int minimize_graph(int32_t precision, const stormm_synthesis *ag_c, stormm_phasespace *ps,
                          const stormm_minimize_controls *mc, stormm_scorecard *sc,
                          stormm_line_min *lm, hipStream_t stream) {
  stormm_synthesis *ag = const_cast<stormm_synthesis *>(ag_c);
  int rc = stormm_phasespace_prime_conjugate_gradient(ps, stream);
  if (rc != STORMM_OK) {
    return rc;
  }
  const int ns = ag->plan.nsys;
  HIP_TRY(hipMemsetAsync(ag->cycles_run.ptr, 0, ns * sizeof(int), stream));
  HIP_TRY(hipMemsetAsync(ag->stop_flags.ptr, 0, ns * sizeof(int), stream));
  HIP_TRY(hipMemsetAsync(ag->rms_force.ptr, 0, ns * sizeof(double), stream));
  const int prec = (precision == STORMM_PREC_DOUBLE) ? 0 : 1;
  MMKit k = make_kit(ag, ps, sc);
  k.store_overwrite = 1;
  set_clash(k, precision, 0.0, 0.0);
  CgKit cg = make_cg_kit(ag, ps, precision);
  cg.steps = ag->cycles_run.as<int>();
  cg.done = ag->stop_flags.as<int>();
  cg.rms = ag->rms_force.as<double>();
  cg.conv_rms = mc->convergence_rms_force;
  cg.max_steps = mc->total_cycles;
  LineKit lk = make_line_kit(ag, ps, sc, lm, precision);
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
  if ((rc = check_mincon(mincon)) != STORMM_OK) {
    return rc;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (persistent == 0) {
    return minimize_stepwise(precision, ag, ps, mincon, sc, lm, st);
  }
  if (persistent == 2) {
    return minimize_graph(precision, ag, ps, mincon, sc, lm, st);
  }
  if (!ag->plan.rst.empty()) {
    return fail(STORMM_ERR_STATE, "the persistent minimizer does not evaluate restraints; "
                                  "minimize a restrained synthesis with the per-step (0) or "
                                  "graph (2) mode");
  }
  MinKit k;
  k.mm = make_kit(ag, ps, sc);
  // The persistent kernel switches to its damped evaluations for cycles below clash_cycles and
  // takes the schedule from its own cycle counter.
  set_clash(k.mm, precision, 0.0, 0.0);
  if (mincon->clash_damping_cycles > 0) {
    k.mm.clash_variant = 1;
    k.mm.clash_cycles = mincon->clash_damping_cycles;
    k.mm.clash_r0 = mincon->clash_minimum_distance;
    k.mm.clash_ratio = mincon->clash_vdw_ratio;
  }
  k.xprv = ps->xprv.as<long long>();
  k.yprv = ps->yprv.as<long long>();
  k.zprv = ps->zprv.as<long long>();
  k.xcg = ps->xcg.as<long long>();
  k.ycg = ps->ycg.as<long long>();
  k.zcg = ps->zcg.as<long long>();
  k.total_cycles = mincon->total_cycles;
  k.sd_cycles = mincon->steepest_descent_cycles;
  k.dx0 = mincon->initial_step;
  k.conv_rms = mincon->convergence_rms_force;
  k.l_move = lm->vals.as<double>();
  k.cycles_run = const_cast<stormm_synthesis *>(ag)->cycles_run.as<int>();
  k.rms_force = const_cast<stormm_synthesis *>(ag)->rms_force.as<double>();
  if ((rc = stormm_scorecard_initialize(sc, stream)) != STORMM_OK) {
    return rc;
  }
  HIP_TRY(launch_minimize_persistent(precision == STORMM_PREC_DOUBLE ? 0 : 1, k, st));
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
  const MMKit k = make_kit(ag, ps, sc);
  const int prec = (precision == STORMM_PREC_DOUBLE) ? 0 : 1;
  HIP_TRY(hipEventRecord(e0, st));
  for (int r = 0; r < repeats; r++) {
    HIP_TRY(launch_mm_eval(prec, k, kTermNonbonded | kTermValence, eval_force, 1, st));
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
