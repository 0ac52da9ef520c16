// Serving pool: a device-resident population of systems that are minimized in continuous
// batches.  Systems are admitted against a device-memory budget, keep every piece of their
// minimization state (fixed-precision coordinates, conjugate-gradient history, line-search
// state, completed cycles, stop flag) in the pool's arrays, and are advanced a few cycles at a
// time together with whatever else is resident.  Each call to stormm_pool_minimize rebuilds the
// work units over the systems still running (work templates are cached per topology), so
// systems can join and leave between calls.  A system's trajectory depends only on its own
// state: alone or among others, in one call or many, it is bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/stormm_amd.h"
#include "host_common.h"
#include "internal.h"
#include "kernels.h"
#include "plan.h"

using namespace stormm_amd;

namespace {

// Device bytes per resident atom: int64 coordinates, forces, previous forces and CG direction
// (12 x 8) plus nine double parameters and the LJ type.
constexpr size_t kPoolBytesPerAtom = (12 * 8) + (9 * 8) + 4;
// Device bytes per system slot: layout (2 ints), step / stop flag (2 ints), RMS force, nine
// line-search doubles and a ScoreCard row.
constexpr size_t kPoolBytesPerSlot = (4 * 4) + 8 + (9 * 8) + (kScoreStride * 8);
// Share of the budget kept for work templates (tile tasks, masks, valence terms, parameters).
constexpr double kTemplateShare = 0.2;

// Upload of a growing host array into a device buffer (reallocated when it grows).
template <typename T>
hipError_t sync_array(DevMem &d, const std::vector<T> &h, size_t &uploaded) {
  if (h.size() == uploaded && d.ptr != nullptr) {
    return hipSuccess;
  }
  const size_t need = std::max<size_t>(h.size(), 1) * sizeof(T);
  if (d.bytes < need) {
    hipError_t e = d.alloc(need + need / 2);
    if (e != hipSuccess) {
      return e;
    }
    uploaded = 0;
  }
  if (h.size() > uploaded) {
    hipError_t e = hipMemcpy(static_cast<char *>(d.ptr) + uploaded * sizeof(T), h.data() + uploaded,
                             (h.size() - uploaded) * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      return e;
    }
  }
  uploaded = h.size();
  return hipSuccess;
}

}  // namespace

struct stormm_pool {
  stormm_pool_config cfg;
  SolventModel sm;
  int cap_atoms = 0, cap_slots = 0;
  size_t template_budget = 0;
  int batch_counter = 0;
  // Host state.
  TemplateStore ts;
  std::map<int, int> free_atoms;        // start -> length (multiples of kGlobalPad)
  std::vector<int> free_slots;
  std::vector<int> slot_start, slot_natom, slot_template, slot_live, slot_table;
  std::vector<int> template_refs;       // resident systems using each template
  // Each admission takes one contiguous atom range, returned when its last system leaves.
  struct Range {
    int start = 0, len = 0, refs = 0;
  };
  std::vector<Range> ranges;
  std::vector<int> free_range_ids;
  std::vector<int> slot_range;          // range of each resident slot
  // Device state.
  DevMem xcrd, ycrd, zcrd, xfrc, yfrc, zfrc, xprv, yprv, zprv, xcg, ycg, zcg;
  DevMem q, ljp1, ljp2, rho, scr, pbr, alpha, beta, gamma, lj_type;
  DevMem starts, natoms, steps, done, rms, line, scorecard, status;
  // Energy history per slot: hist_len ScoreCard rows, one every cfg.minimize.
  // diagnostic_print_frequency cycles (none when that is 0).
  int hist_len = 0;
  DevMem hist;
  DevMem tasks, masks, vterms, bond_par, angl_par, dihe_par, dihe14, ljab, ljab14;
  size_t up_tasks = 0, up_masks = 0, up_vterms = 0, up_bond = 0, up_angl = 0, up_dihe = 0,
         up_d14 = 0, up_ljab = 0, up_ljab14 = 0;
  DevMem wu, wu_sys, active, init_buf;
};

static int check_config(const stormm_pool_config *c) {
  if (c->precision != STORMM_PREC_SINGLE && c->precision != STORMM_PREC_DOUBLE) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "pool precision must be single or double");
  }
  if (c->gpos_bits < 24 || c->gpos_bits > 38 || c->frc_bits < 18 || c->frc_bits > 40 ||
      c->nrg_bits < 22 || c->nrg_bits > 40) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "fixed-precision bits out of range");
  }
  if (c->max_systems <= 0 || c->hbm_budget_bytes <= 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "pool needs a positive budget and slot count");
  }
  if (c->minimize.total_cycles < 0 || c->minimize.steepest_descent_cycles < 0 ||
      !(c->minimize.initial_step > 0.0) || c->minimize.clash_damping_cycles < 0 ||
      c->minimize.clash_damping_cycles > c->minimize.total_cycles ||
      c->minimize.clash_minimum_distance < 0.0 || c->minimize.clash_vdw_ratio < 0.0 ||
      c->minimize.diagnostic_print_frequency < 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid minimization controls");
  }
  if (c->minimize.diagnostic_print_frequency > 0 && c->minimize.total_cycles == 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "an energy history needs a cycle limit");
  }
  return STORMM_OK;
}

// Samples of each slot's energy history.
static int history_length(const stormm_pool_config *c) {
  const int every = c->minimize.diagnostic_print_frequency;
  return (every > 0) ? (c->minimize.total_cycles + every - 1) / every : 0;
}

extern "C" int stormm_pool_create(const stormm_pool_config *cfg, stormm_pool **out) {
  if (cfg == nullptr || out == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  int rc = check_config(cfg);
  if (rc != STORMM_OK) {
    return rc;
  }
  std::unique_ptr<stormm_pool> p(new stormm_pool());
  p->cfg = *cfg;
  stormm_topology_batch probe;
  std::memset(&probe, 0, sizeof(probe));
  probe.igb = cfg->igb;
  probe.dielectric = cfg->dielectric;
  probe.salt_concentration = cfg->salt_concentration;
  probe.coulomb_constant = cfg->coulomb_constant;
  rc = solvent_model(&probe, p->sm);
  if (rc != STORMM_OK) {
    return rc;
  }
  const size_t budget = static_cast<size_t>(cfg->hbm_budget_bytes);
  p->hist_len = history_length(cfg);
  const size_t hist_bytes = static_cast<size_t>(p->hist_len) * kScoreStride * sizeof(long long);
  const size_t slot_bytes = static_cast<size_t>(cfg->max_systems) *
                            (kPoolBytesPerSlot + hist_bytes);
  p->template_budget = static_cast<size_t>(kTemplateShare * static_cast<double>(budget));
  if (budget <= slot_bytes + p->template_budget) {
    return fail(STORMM_ERR_CAPACITY, "budget too small for the requested system slots");
  }
  const size_t atoms = (budget - slot_bytes - p->template_budget) / kPoolBytesPerAtom;
  p->cap_atoms = static_cast<int>(std::min<size_t>(atoms / kGlobalPad * kGlobalPad,
                                                   static_cast<size_t>(1) << 30));
  p->cap_slots = cfg->max_systems;
  if (p->cap_atoms < kGlobalPad) {
    return fail(STORMM_ERR_CAPACITY, "budget holds no atoms");
  }
  const size_t na = static_cast<size_t>(p->cap_atoms), ns = static_cast<size_t>(p->cap_slots);
  for (DevMem *d : {&p->xcrd, &p->ycrd, &p->zcrd, &p->xfrc, &p->yfrc, &p->zfrc, &p->xprv,
                    &p->yprv, &p->zprv, &p->xcg, &p->ycg, &p->zcg}) {
    HIP_TRY(d->alloc(na * sizeof(long long)));
    HIP_TRY(hipMemset(d->ptr, 0, na * sizeof(long long)));
  }
  for (DevMem *d : {&p->q, &p->ljp1, &p->ljp2, &p->rho, &p->scr, &p->pbr, &p->alpha, &p->beta,
                    &p->gamma}) {
    HIP_TRY(d->alloc(na * sizeof(double)));
  }
  HIP_TRY(p->lj_type.alloc(na * sizeof(int)));
  HIP_TRY(p->starts.alloc(ns * sizeof(int)));
  HIP_TRY(p->natoms.alloc(ns * sizeof(int)));
  HIP_TRY(p->steps.alloc(ns * sizeof(int)));
  HIP_TRY(p->done.alloc(ns * sizeof(int)));
  HIP_TRY(p->rms.alloc(ns * sizeof(double)));
  HIP_TRY(p->line.alloc(9 * ns * sizeof(double)));
  HIP_TRY(p->scorecard.alloc(ns * kScoreStride * sizeof(long long)));
  HIP_TRY(hipMemset(p->scorecard.ptr, 0, ns * kScoreStride * sizeof(long long)));
  if (p->hist_len > 0) {
    HIP_TRY(p->hist.alloc(ns * hist_bytes));
  }
  HIP_TRY(p->status.alloc(sizeof(int)));
  HIP_TRY(hipMemset(p->status.ptr, 0, sizeof(int)));
  HIP_TRY(p->active.alloc(ns * sizeof(int)));
  p->free_atoms[0] = p->cap_atoms;
  for (int s = p->cap_slots - 1; s >= 0; s--) {
    p->free_slots.push_back(s);
  }
  p->slot_start.assign(ns, 0);
  p->slot_natom.assign(ns, 0);
  p->slot_template.assign(ns, -1);
  p->slot_live.assign(ns, 0);
  p->slot_table.assign(ns, 0);
  p->slot_range.assign(ns, -1);
  *out = p.release();
  return STORMM_OK;
}

extern "C" int stormm_pool_destroy(stormm_pool *pool) {
  delete pool;
  return STORMM_OK;
}

extern "C" int stormm_pool_capacity(const stormm_pool *pool, int64_t *free_atoms,
                                    int32_t *free_slots, int64_t *template_bytes_used,
                                    int64_t *capacity_atoms) {
  if (pool == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null pool");
  }
  long long fa = 0;
  for (const auto &r : pool->free_atoms) {
    fa += r.second;
  }
  if (free_atoms) {
    *free_atoms = fa;
  }
  if (free_slots) {
    *free_slots = static_cast<int32_t>(pool->free_slots.size());
  }
  if (template_bytes_used) {
    *template_bytes_used = static_cast<int64_t>(template_store_bytes(pool->ts));
  }
  if (capacity_atoms) {
    *capacity_atoms = pool->cap_atoms;
  }
  return STORMM_OK;
}

// First-fit allocation of `len` atoms; returns the start or -1.
static int take_atoms(stormm_pool *p, int len) {
  for (auto it = p->free_atoms.begin(); it != p->free_atoms.end(); ++it) {
    if (it->second >= len) {
      const int start = it->first;
      const int rest = it->second - len;
      p->free_atoms.erase(it);
      if (rest > 0) {
        p->free_atoms[start + len] = rest;
      }
      return start;
    }
  }
  return -1;
}

static void give_atoms(stormm_pool *p, int start, int len) {
  auto it = p->free_atoms.emplace(start, len).first;
  auto next = std::next(it);
  if (next != p->free_atoms.end() && it->first + it->second == next->first) {
    it->second += next->second;
    p->free_atoms.erase(next);
  }
  if (it != p->free_atoms.begin()) {
    auto prev = std::prev(it);
    if (prev->first + prev->second == it->first) {
      prev->second += it->second;
      p->free_atoms.erase(it);
    }
  }
}

// Uploads the work templates admitted since the last upload (device arrays grow as needed).
static hipError_t sync_templates(stormm_pool *p) {
  const hipError_t steps[] = {
      sync_array(p->tasks, p->ts.tasks, p->up_tasks),
      sync_array(p->masks, p->ts.masks, p->up_masks),
      sync_array(p->vterms, p->ts.vterms, p->up_vterms),
      sync_array(p->bond_par, p->ts.bond_par, p->up_bond),
      sync_array(p->angl_par, p->ts.angl_par, p->up_angl),
      sync_array(p->dihe_par, p->ts.dihe_par, p->up_dihe),
      sync_array(p->dihe14, p->ts.dihe14, p->up_d14),
      sync_array(p->ljab, p->ts.ljab, p->up_ljab),
      sync_array(p->ljab14, p->ts.ljab14, p->up_ljab14)};
  for (hipError_t e : steps) {
    if (e != hipSuccess) {
      return e;
    }
  }
  return hipSuccess;
}

// Whether every template array on the device matches the host store.
static bool templates_in_sync(const stormm_pool *p) {
  return p->tasks.ptr != nullptr && p->up_tasks == p->ts.tasks.size() &&
         p->up_masks == p->ts.masks.size() && p->up_vterms == p->ts.vterms.size() &&
         p->up_bond == p->ts.bond_par.size() && p->up_angl == p->ts.angl_par.size() &&
         p->up_dihe == p->ts.dihe_par.size() && p->up_d14 == p->ts.dihe14.size() &&
         p->up_ljab == p->ts.ljab.size() && p->up_ljab14 == p->ts.ljab14.size();
}

// The next upload rewrites every template array from the start.
static void invalidate_template_upload(stormm_pool *p) {
  p->up_tasks = p->up_masks = p->up_vterms = p->up_bond = p->up_angl = p->up_dihe = 0;
  p->up_d14 = p->up_ljab = p->up_ljab14 = 0;
}

// Drops the templates that no resident system uses and renumbers the rest; false when every
// template is in use.
static bool drop_unused_templates(stormm_pool *p) {
  std::vector<char> keep(p->ts.templates.size());
  bool any_dropped = false;
  for (size_t t = 0; t < keep.size(); t++) {
    keep[t] = p->template_refs[t] > 0;
    any_dropped = any_dropped || !keep[t];
  }
  if (!any_dropped) {
    return false;
  }
  std::vector<int> remap;
  compact_templates(p->ts, keep, remap);
  std::vector<int> refs(p->ts.templates.size(), 0);
  for (size_t t = 0; t < remap.size(); t++) {
    if (remap[t] >= 0) {
      refs[remap[t]] = p->template_refs[t];
    }
  }
  p->template_refs.swap(refs);
  for (int s = 0; s < p->cap_slots; s++) {
    p->slot_template[s] = p->slot_live[s] ? remap[p->slot_template[s]] : -1;
  }
  invalidate_template_upload(p);
  return true;
}

// Templates of the admitted entries (tids[e]), built or found in the store.  Nothing is kept on
// failure; when the new templates would exceed the template budget, the templates no resident
// system uses are dropped and the search runs once more.  `mark` receives the store's state
// before this admission's templates.
static int admit_templates(stormm_pool *p, const stormm_topology_batch *batch,
                           const std::vector<SystemTerms> &st, const std::vector<int> &sys_of,
                           int batch_id, std::vector<int> &tids, TemplateStoreMark &mark) {
  for (int attempt = 0; attempt < 2; attempt++) {
    mark = mark_templates(p->ts);
    std::vector<int> tid_of_sys(batch->system_count, -1);
    int rc = STORMM_OK;
    for (size_t e = 0; e < sys_of.size() && rc == STORMM_OK; e++) {
      const int s = sys_of[e];
      if (tid_of_sys[s] < 0) {
        rc = find_or_add_template(batch, st, s, batch_id, p->ts, &tid_of_sys[s]);
      }
      tids[e] = tid_of_sys[s];
    }
    if (rc != STORMM_OK) {
      rollback_templates(p->ts, mark);
      return rc;
    }
    if (template_store_bytes(p->ts) <= p->template_budget) {
      p->template_refs.resize(p->ts.templates.size(), 0);
      return STORMM_OK;
    }
    rollback_templates(p->ts, mark);
    if (attempt > 0 || !drop_unused_templates(p)) {
      break;
    }
  }
  return fail(STORMM_ERR_CAPACITY, "work templates exceed the pool's template budget");
}

// Admits `count` systems; entry e is system sys_of[e] of the batch with coordinates at
// xyz + 3 * (first atom of entry e), entries laid out one after another.
static int admit_entries(stormm_pool *pool, const stormm_topology_batch *batch,
                         const std::vector<int> &sys_of, const double *xyz, int32_t *slots) {
  if (batch->igb != pool->cfg.igb) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "the batch's implicit-solvent model differs from "
                                             "the pool's");
  }
  if (batch->restraint_counts != nullptr) {
    for (int s = 0; s < batch->system_count; s++) {
      if (batch->restraint_counts[s] != 0) {
        return fail(STORMM_ERR_INVALID_ARGUMENT, "the pool does not take restrained systems "
                                                 "(minimize them with a synthesis)");
      }
    }
  }
  stormm_pool *p = pool;
  const int ns = static_cast<int>(sys_of.size());
  if (static_cast<int>(p->free_slots.size()) < ns) {
    return fail(STORMM_ERR_CAPACITY, "no free system slots");
  }
  std::vector<SystemTerms> st;
  int rc = collect_terms(batch, st);
  if (rc != STORMM_OK) {
    return rc;
  }
  int total = 0;
  for (int e = 0; e < ns; e++) {
    total += round_up(batch->atom_counts[sys_of[e]], kGlobalPad);
  }
  // Templates first (they may be shared with resident systems), within the template budget.
  const int batch_id = ++p->batch_counter;
  std::vector<int> tids(ns);
  TemplateStoreMark mark;
  rc = admit_templates(p, batch, st, sys_of, batch_id, tids, mark);
  if (rc != STORMM_OK) {
    return rc;
  }
  for (TopoTemplate &tp : p->ts.templates) {
    if (tp.batch_id == batch_id) {
      tp.batch_id = -1;  // later batches match it by hash and counts
    }
  }
  const int start = take_atoms(p, total);
  if (start < 0) {
    rollback_templates(p->ts, mark);
    return fail(STORMM_ERR_CAPACITY, "not enough free atoms in the pool's budget");
  }
  // Parameters, coordinates and fresh minimization state of the new systems, staged for one
  // contiguous range.
  AtomParams ap;
  ap.assign(static_cast<size_t>(total));
  std::vector<long long> cx(total, 0), cy(total, 0), cz(total, 0);
  const double scale = std::ldexp(1.0, p->cfg.gpos_bits);
  int off = 0;
  size_t xoff = 0;
  for (int e = 0; e < ns; e++) {
    const int s = sys_of[e];
    const int n = batch->atom_counts[s];
    rc = fill_atom_parameters(batch, st[s], n, p->ts.templates[tids[e]], p->sm, off, ap, s);
    if (rc != STORMM_OK) {
      give_atoms(p, start, total);
      rollback_templates(p->ts, mark);
      return rc;
    }
    for (int i = 0; i < n; i++) {
      const double *r = xyz + 3 * (xoff + static_cast<size_t>(i));
      cx[off + i] = static_cast<long long>(std::llround(r[0] * scale));
      cy[off + i] = static_cast<long long>(std::llround(r[1] * scale));
      cz[off + i] = static_cast<long long>(std::llround(r[2] * scale));
    }
    off += round_up(n, kGlobalPad);
    xoff += static_cast<size_t>(n);
  }
  // Work templates reach the device with their first admission, so that minimization calls
  // never allocate for them.
  {
    const hipError_t e = sync_templates(p);
    if (e != hipSuccess) {
      give_atoms(p, start, total);
      rollback_templates(p->ts, mark);
      invalidate_template_upload(p);
      return fail(STORMM_ERR_HIP, std::string("template upload: ") + hipGetErrorString(e));
    }
  }
  auto put = [&](DevMem &d, const void *src, size_t elt) {
    return hipMemcpy(static_cast<char *>(d.ptr) + static_cast<size_t>(start) * elt, src,
                     static_cast<size_t>(total) * elt, hipMemcpyHostToDevice);
  };
  auto zero = [&](DevMem &d) {
    return hipMemset(static_cast<char *>(d.ptr) + static_cast<size_t>(start) * sizeof(long long),
                     0, static_cast<size_t>(total) * sizeof(long long));
  };
  HIP_TRY(put(p->xcrd, cx.data(), sizeof(long long)));
  HIP_TRY(put(p->ycrd, cy.data(), sizeof(long long)));
  HIP_TRY(put(p->zcrd, cz.data(), sizeof(long long)));
  for (DevMem *d : {&p->xfrc, &p->yfrc, &p->zfrc, &p->xprv, &p->yprv, &p->zprv, &p->xcg,
                    &p->ycg, &p->zcg}) {
    HIP_TRY(zero(*d));
  }
  HIP_TRY(put(p->q, ap.q.data(), sizeof(double)));
  HIP_TRY(put(p->ljp1, ap.ljp1.data(), sizeof(double)));
  HIP_TRY(put(p->ljp2, ap.ljp2.data(), sizeof(double)));
  HIP_TRY(put(p->rho, ap.rho.data(), sizeof(double)));
  HIP_TRY(put(p->scr, ap.scr.data(), sizeof(double)));
  HIP_TRY(put(p->pbr, ap.pbr.data(), sizeof(double)));
  HIP_TRY(put(p->alpha, ap.alpha.data(), sizeof(double)));
  HIP_TRY(put(p->beta, ap.beta.data(), sizeof(double)));
  HIP_TRY(put(p->gamma, ap.gamma.data(), sizeof(double)));
  HIP_TRY(put(p->lj_type, ap.lj_type.data(), sizeof(int)));
  off = 0;
  int rid;
  if (!p->free_range_ids.empty()) {
    rid = p->free_range_ids.back();
    p->free_range_ids.pop_back();
  }
  else {
    rid = static_cast<int>(p->ranges.size());
    p->ranges.emplace_back();
  }
  p->ranges[rid].start = start;
  p->ranges[rid].len = total;
  p->ranges[rid].refs = ns;
  std::vector<int> e_slot(ns), e_start(ns), e_natom(ns);
  for (int e = 0; e < ns; e++) {
    const int s = sys_of[e];
    const int slot = p->free_slots.back();
    p->free_slots.pop_back();
    p->slot_start[slot] = start + off;
    p->slot_natom[slot] = batch->atom_counts[s];
    p->slot_template[slot] = tids[e];
    p->slot_live[slot] = 1;
    p->slot_table[slot] = p->ts.templates[tids[e]].lb ? 0 : 1;
    p->slot_range[slot] = rid;
    p->template_refs[tids[e]]++;
    off += round_up(batch->atom_counts[s], kGlobalPad);
    e_slot[e] = slot;
    e_start[e] = p->slot_start[slot];
    e_natom[e] = p->slot_natom[slot];
    slots[e] = slot;
  }
  // Fresh per-system state (LineMinimization::primeMoveLengths: l_move = dx0, the rest zero).
  if (p->init_buf.bytes < 3 * static_cast<size_t>(ns) * sizeof(int)) {
    HIP_TRY(p->init_buf.alloc(6 * static_cast<size_t>(ns) * sizeof(int)));
  }
  int *ib = p->init_buf.as<int>();
  HIP_TRY(hipMemcpy(ib, e_slot.data(), ns * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ib + ns, e_start.data(), ns * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ib + 2 * ns, e_natom.data(), ns * sizeof(int), hipMemcpyHostToDevice));
  SlotInitKit ik;
  ik.count = ns;
  ik.cap_slots = p->cap_slots;
  ik.slots = ib;
  ik.starts = ib + ns;
  ik.natoms = ib + 2 * ns;
  ik.slot_start = p->starts.as<int>();
  ik.slot_natom = p->natoms.as<int>();
  ik.steps = p->steps.as<int>();
  ik.done = p->done.as<int>();
  ik.rms = p->rms.as<double>();
  ik.line = p->line.as<double>();
  ik.scorecard = p->scorecard.as<long long>();
  ik.dx0 = p->cfg.minimize.initial_step;
  HIP_TRY(launch_pool_init_slots(ik, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  return STORMM_OK;
}

extern "C" int stormm_pool_admit(stormm_pool *pool, const stormm_topology_batch *batch,
                                 const double *xyz, int32_t *slots) {
  if (pool == nullptr || batch == nullptr || xyz == nullptr || slots == nullptr ||
      batch->system_count <= 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid admission request");
  }
  std::vector<int> sys_of(batch->system_count);
  for (int s = 0; s < batch->system_count; s++) {
    sys_of[s] = s;
  }
  return admit_entries(pool, batch, sys_of, xyz, slots);
}

extern "C" int stormm_pool_admit_conformers(stormm_pool *pool,
                                            const stormm_topology_batch *molecule, int32_t count,
                                            const double *xyz, int32_t *slots) {
  if (pool == nullptr || molecule == nullptr || xyz == nullptr || slots == nullptr ||
      molecule->system_count != 1 || count <= 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid conformer admission request (the "
                                             "topology batch must hold one system)");
  }
  return admit_entries(pool, molecule, std::vector<int>(count, 0), xyz, slots);
}

extern "C" int stormm_pool_release(stormm_pool *pool, const int32_t *slots, int32_t count) {
  if (pool == nullptr || (count > 0 && slots == nullptr)) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid release request");
  }
  std::vector<char> seen(pool->cap_slots, 0);
  for (int i = 0; i < count; i++) {
    const int s = slots[i];
    if (s < 0 || s >= pool->cap_slots || !pool->slot_live[s] || seen[s]) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "slot " + std::to_string(s) +
                                                   " is not resident (or listed twice)");
    }
    seen[s] = 1;
  }
  for (int i = 0; i < count; i++) {
    const int s = slots[i];
    pool->slot_live[s] = 0;
    pool->template_refs[pool->slot_template[s]]--;
    pool->slot_template[s] = -1;
    const int rid = pool->slot_range[s];
    pool->slot_range[s] = -1;
    if (--pool->ranges[rid].refs == 0) {
      give_atoms(pool, pool->ranges[rid].start, pool->ranges[rid].len);
      pool->free_range_ids.push_back(rid);
    }
    pool->free_slots.push_back(s);
  }
  return STORMM_OK;
}

// Kits over the pool's arrays for the work units of one call (wu, wu_sys: device arrays).
static MMKit pool_mm_kit(stormm_pool *p, int nwu, int maxa, int block, const WuDesc *wu,
                         const WuSys *wu_sys) {
  MMKit k;
  std::memset(&k, 0, sizeof(k));
  k.nsys = p->cap_slots;
  k.nwu = nwu;
  k.maxa = round_up(std::max(maxa, kTile), kTile);
  k.block = block;
  k.gpos_bits = p->cfg.gpos_bits;
  k.frc_bits = p->cfg.frc_bits;
  k.nrg_bits = p->cfg.nrg_bits;
  k.igb = p->sm.igb;
  k.dielectric = p->sm.dielectric;
  k.kappa = p->sm.kappa;
  k.sys_start = p->starts.as<int>();
  k.sys_natom = p->natoms.as<int>();
  k.wu = wu;
  k.wu_sys = wu_sys;
  k.tasks = p->tasks.as<Task>();
  k.masks = p->masks.as<unsigned>();
  k.vterms = p->vterms.as<int4>();
  k.bond_par = p->bond_par.as<double2>();
  k.angl_par = p->angl_par.as<double2>();
  k.dihe_par = p->dihe_par.as<DiheParam>();
  k.dihe14 = p->dihe14.as<double2>();
  k.q = p->q.as<double>();
  k.ljp1 = p->ljp1.as<double>();
  k.ljp2 = p->ljp2.as<double>();
  k.lj_type = p->lj_type.as<int>();
  k.ljab = p->ljab.as<double2>();
  k.ljab14 = p->ljab14.as<double2>();
  k.rho = p->rho.as<double>();
  k.scr = p->scr.as<double>();
  k.pbr = p->pbr.as<double>();
  k.gb_alpha = p->alpha.as<double>();
  k.gb_beta = p->beta.as<double>();
  k.gb_gamma = p->gamma.as<double>();
  k.xcrd = p->xcrd.as<long long>();
  k.ycrd = p->ycrd.as<long long>();
  k.zcrd = p->zcrd.as<long long>();
  k.xfrc = p->xfrc.as<long long>();
  k.yfrc = p->yfrc.as<long long>();
  k.zfrc = p->zfrc.as<long long>();
  k.scorecard = p->scorecard.as<long long>();
  k.status = p->status.as<int>();
  k.store_overwrite = 1;
  return k;
}

// Where one call's work units, their system entries and the list of running slots go: the
// pool's own (growing) buffers, or a caller-provided region of fixed size.
struct PoolStaging {
  void *base = nullptr;   // null: the pool's buffers
  size_t bytes = 0;
};

static size_t align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

static int pool_minimize(stormm_pool *pool, int32_t cycles, int32_t *running, void *stream,
                         const PoolStaging &staging) {
  if (pool == nullptr || cycles < 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid minimization request");
  }
  stormm_pool *p = pool;
  hipStream_t st = static_cast<hipStream_t>(stream);
  // Systems still running: resident, not stopped, under the cycle limit.
  std::vector<int> steps(p->cap_slots), done(p->cap_slots);
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipMemcpy(steps.data(), p->steps.ptr, p->cap_slots * sizeof(int),
                    hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(done.data(), p->done.ptr, p->cap_slots * sizeof(int), hipMemcpyDeviceToHost));
  const int max_steps = p->cfg.minimize.total_cycles;
  std::vector<int> act;
  int min_step = max_steps;
  for (int s = 0; s < p->cap_slots; s++) {
    if (p->slot_live[s] && !done[s] && (max_steps == 0 || steps[s] < max_steps)) {
      act.push_back(s);
      min_step = std::min(min_step, steps[s]);
    }
  }
  if (running) {
    *running = static_cast<int32_t>(act.size());
  }
  if (act.empty() || cycles == 0) {
    return STORMM_OK;
  }
  // The admissions uploaded every template this call uses.
  if (!templates_in_sync(p)) {
    return fail(STORMM_ERR_STATE, "the pool's work templates are not on the device");
  }
  const int na = static_cast<int>(act.size());
  std::vector<double> cost(na);
  std::vector<int> natom(na), start(na), tmpl(na), index(na);
  std::vector<char> table(na);
  for (int i = 0; i < na; i++) {
    const int s = act[i];
    cost[i] = p->ts.templates[p->slot_template[s]].cost;
    natom[i] = p->slot_natom[s];
    start[i] = p->slot_start[s];
    tmpl[i] = p->slot_template[s];
    table[i] = static_cast<char>(p->slot_table[s]);
    index[i] = s;
  }
  const stormm_launch_shape &sh = p->cfg.shape;
  const int cap = (sh.wu_atom_capacity > 0) ? sh.wu_atom_capacity : kMaxWuAtoms;
  const int per_cu = (sh.target_wu_per_cu > 0) ? sh.target_wu_per_cu : 2;
  const int block = (sh.block_threads > 0) ? sh.block_threads : kEvalBlock;
  std::vector<std::vector<int>> bins;
  std::vector<int> modes;
  // The resident set changes every call; balancing units one by one (largest first) serves
  // these short launches better than the compute-unit groups of the batch synthesis.
  pack_systems(cost, natom, table, cap, cu_count_or_default(), per_cu, false, bins, modes);
  std::vector<WuDesc> wu;
  std::vector<WuSys> wu_sys;
  int maxa = 0;
  for (size_t u = 0; u < bins.size(); u++) {
    maxa = std::max(maxa, append_work_unit(bins[u], modes[u], index, start, natom, tmpl, p->ts,
                                           wu, wu_sys));
    if (wu.back().task_count > kMaxWuTasks) {
      return fail(STORMM_ERR_CAPACITY, "a work unit holds more than " +
                                           std::to_string(kMaxWuTasks) + " tile tasks");
    }
  }
  const size_t wu_bytes = wu.size() * sizeof(WuDesc);
  const size_t wu_sys_bytes = wu_sys.size() * sizeof(WuSys);
  const size_t act_bytes = act.size() * sizeof(int);
  WuDesc *d_wu;
  WuSys *d_wu_sys;
  int *d_active;
  if (staging.base == nullptr) {
    if (p->wu.bytes < wu_bytes) {
      HIP_TRY(p->wu.alloc(2 * wu_bytes));
    }
    if (p->wu_sys.bytes < wu_sys_bytes) {
      HIP_TRY(p->wu_sys.alloc(2 * wu_sys_bytes));
    }
    d_wu = p->wu.as<WuDesc>();
    d_wu_sys = p->wu_sys.as<WuSys>();
    d_active = p->active.as<int>();
  }
  else {
    const size_t need = align256(wu_bytes) + align256(wu_sys_bytes) + align256(act_bytes);
    if (need > staging.bytes) {
      return fail(STORMM_ERR_CAPACITY,
                  "the call's work units need " + std::to_string(need) + " bytes of staging (" +
                      std::to_string(wu.size()) + " units, " + std::to_string(na) +
                      " systems); the allocation-free entry point holds " +
                      std::to_string(staging.bytes));
    }
    char *b = static_cast<char *>(staging.base);
    d_wu = reinterpret_cast<WuDesc *>(b);
    d_wu_sys = reinterpret_cast<WuSys *>(b + align256(wu_bytes));
    d_active = reinterpret_cast<int *>(b + align256(wu_bytes) + align256(wu_sys_bytes));
  }
  HIP_TRY(hipMemcpy(d_wu, wu.data(), wu_bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_wu_sys, wu_sys.data(), wu_sys_bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_active, act.data(), act_bytes, hipMemcpyHostToDevice));
  const int prec = (p->cfg.precision == STORMM_PREC_DOUBLE) ? 0 : 1;
  const MMKit mm = pool_mm_kit(p, static_cast<int>(wu.size()), maxa, block, d_wu, d_wu_sys);
  CgKit cg;
  std::memset(&cg, 0, sizeof(cg));
  cg.nsys = na;
  cg.slots = d_active;
  cg.sys_start = p->starts.as<int>();
  cg.sys_natom = p->natoms.as<int>();
  cg.steps = p->steps.as<int>();
  cg.done = p->done.as<int>();
  cg.rms = p->rms.as<double>();
  cg.conv_rms = p->cfg.minimize.convergence_rms_force;
  cg.max_steps = max_steps;
  cg.xfrc = mm.xfrc;
  cg.yfrc = mm.yfrc;
  cg.zfrc = mm.zfrc;
  cg.xprv = p->xprv.as<long long>();
  cg.yprv = p->yprv.as<long long>();
  cg.zprv = p->zprv.as<long long>();
  cg.xcg = p->xcg.as<long long>();
  cg.ycg = p->ycg.as<long long>();
  cg.zcg = p->zcg.as<long long>();
  cg.inv_frc_scale = std::ldexp(1.0, -p->cfg.frc_bits);
  cg.dp_mode = (prec == 0) ? 1 : 0;
  if (p->hist_len > 0) {
    // Each slot's sample k is its energy at the start of cycle k * every, counted from its
    // admission.
    HistKit &h = cg.hist;
    h.rows = p->hist.as<long long>();
    h.scorecard = mm.scorecard;
    h.sample_stride = kScoreStride;
    h.sys_stride = static_cast<long long>(p->hist_len) * kScoreStride;
    h.every = p->cfg.minimize.diagnostic_print_frequency;
    h.first = 0;
    h.capacity = p->hist_len;
    h.max_cycle = max_steps;
    h.inv_nrg_scale = std::ldexp(1.0, -p->cfg.nrg_bits);
  }
  LineKit lk;
  std::memset(&lk, 0, sizeof(lk));
  lk.nsys = na;
  lk.slots = d_active;
  lk.done = p->done.as<int>();
  lk.sys_start = cg.sys_start;
  lk.sys_natom = cg.sys_natom;
  lk.xcrd = mm.xcrd;
  lk.ycrd = mm.ycrd;
  lk.zcrd = mm.zcrd;
  lk.xfrc = mm.xfrc;
  lk.yfrc = mm.yfrc;
  lk.zfrc = mm.zfrc;
  lk.scorecard = mm.scorecard;
  lk.inv_nrg_scale = std::ldexp(1.0, -p->cfg.nrg_bits);
  lk.inv_frc_scale = cg.inv_frc_scale;
  lk.gpos_scale = std::ldexp(1.0, p->cfg.gpos_bits);
  double *v = p->line.as<double>();
  const size_t n = static_cast<size_t>(p->cap_slots);
  lk.l_move = v;
  lk.s_move = v + n;
  lk.mfac_a = v + 2 * n;
  lk.mfac_b = v + 3 * n;
  lk.mfac_c = v + 4 * n;
  lk.nrg_a = v + 5 * n;
  lk.nrg_b = v + 6 * n;
  lk.nrg_c = v + 7 * n;
  lk.nrg_d = v + 8 * n;
  lk.dp_mode = cg.dp_mode;
  const int terms = kTermNonbonded | kTermValence;
  const int sd = p->cfg.minimize.steepest_descent_cycles;
  // Clash damping follows each system's own cycle count (hpc_minimization.cu:269-281): the
  // damped kernels read it from the device (before the conjugate-gradient kernel advances it
  // for the force evaluation, after for the energy evaluations) and give the plain potentials
  // to systems past their damping cycles.  They run while the least advanced system may still
  // be damped.
  const int cdcyc = p->cfg.minimize.clash_damping_cycles;
  MMKit mf_clash = mm, me_clash = mm;
  mf_clash.clash_variant = 1;
  mf_clash.clash_cycles = cdcyc;
  mf_clash.clash_step = p->steps.as<int>();
  mf_clash.clash_r0 = p->cfg.minimize.clash_minimum_distance;
  mf_clash.clash_ratio = p->cfg.minimize.clash_vdw_ratio;
  me_clash = mf_clash;
  me_clash.clash_step_offset = -1;
  // The reference's per-cycle sequence (src/MolecularMechanics/hpc_minimization.cu:262-405):
  // force + energy evaluation, conjugate-gradient direction, then four line moves separated by
  // energy evaluations.  Stores overwrite, so no force or ScoreCard zeroing is launched.
  for (int c = 0; c < cycles; c++) {
    const bool damp = (min_step + c < cdcyc);
    HIP_TRY(launch_mm_eval(prec, damp ? mf_clash : mm, terms, 1, 1, st));
    HIP_TRY(launch_conjugate_gradient(cg, 0, sd, st));
    for (int move = 0; move < 4; move++) {
      if (move > 0) {
        HIP_TRY(launch_mm_eval(prec, damp ? me_clash : mm, terms, 0, 1, st));
      }
      HIP_TRY(launch_line_advance(lk, move, st));
    }
  }
  // Energies at the coordinates the call leaves.
  HIP_TRY(launch_mm_eval(prec, mm, terms, 0, 1, st));
  if (staging.base != nullptr) {
    // The staging region is shared by every call on the device: it must not be rewritten
    // before this call's kernels have read it.
    HIP_TRY(hipStreamSynchronize(st));
  }
  return STORMM_OK;
}

extern "C" int stormm_pool_minimize(stormm_pool *pool, int32_t cycles, int32_t *running,
                                    void *stream) {
  return pool_minimize(pool, cycles, running, stream, PoolStaging());
}

extern "C" int stormm_pool_status(const stormm_pool *pool, const int32_t *slots, int32_t count,
                                  int32_t *steps, int32_t *done, double *rms, double *energy) {
  if (pool == nullptr || (count > 0 && slots == nullptr)) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid status request");
  }
  const stormm_pool *p = pool;
  HIP_TRY(hipDeviceSynchronize());
  std::vector<int> hs(p->cap_slots), hd(p->cap_slots);
  std::vector<double> hr(p->cap_slots);
  std::vector<long long> sc(static_cast<size_t>(p->cap_slots) * kScoreStride);
  HIP_TRY(hipMemcpy(hs.data(), p->steps.ptr, hs.size() * sizeof(int), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(hd.data(), p->done.ptr, hd.size() * sizeof(int), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(hr.data(), p->rms.ptr, hr.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(sc.data(), p->scorecard.ptr, sc.size() * sizeof(long long),
                    hipMemcpyDeviceToHost));
  const double inv_nrg = std::ldexp(1.0, -p->cfg.nrg_bits);
  for (int i = 0; i < count; i++) {
    const int s = slots[i];
    if (s < 0 || s >= p->cap_slots || !p->slot_live[s]) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "slot " + std::to_string(s) + " is not resident");
    }
    const int max_steps = p->cfg.minimize.total_cycles;
    if (steps) {
      steps[i] = hs[s];
    }
    if (done) {
      done[i] = (hd[s] != 0 || (max_steps > 0 && hs[s] >= max_steps)) ? 1 : 0;
    }
    if (rms) {
      rms[i] = hr[s];
    }
    if (energy) {
      long long acc = 0;
      for (int c = 0; c < kNrgSlots; c++) {
        acc += sc[static_cast<size_t>(s) * kScoreStride + c];
      }
      energy[i] = static_cast<double>(acc) * inv_nrg;
    }
  }
  return STORMM_OK;
}

extern "C" int stormm_pool_history(const stormm_pool *pool, int32_t slot, int32_t capacity,
                                   int32_t *count, double *energy, int64_t *rows) {
  if (pool == nullptr || count == nullptr || capacity < 0 || slot < 0 ||
      slot >= pool->cap_slots || !pool->slot_live[slot]) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "invalid history request");
  }
  const stormm_pool *p = pool;
  *count = 0;
  if (p->hist_len == 0) {
    return STORMM_OK;
  }
  HIP_TRY(hipDeviceSynchronize());
  int steps = 0, done = 0;
  HIP_TRY(hipMemcpy(&steps, p->steps.as<int>() + slot, sizeof(int), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&done, p->done.as<int>() + slot, sizeof(int), hipMemcpyDeviceToHost));
  // Snapshots of the cycles run (0, every, ... below steps) and of the cycle at which the RMS
  // force stopped the system.
  const int every = p->cfg.minimize.diagnostic_print_frequency;
  const int max_steps = p->cfg.minimize.total_cycles;
  int n = (steps + every - 1) / every;
  if (done != 0 && steps % every == 0 && steps < max_steps) {
    n += 1;
  }
  n = std::min(n, p->hist_len);
  if (n > capacity) {
    return fail(STORMM_ERR_CAPACITY, "slot " + std::to_string(slot) + " holds " +
                                         std::to_string(n) + " samples; room for " +
                                         std::to_string(capacity) + " given");
  }
  std::vector<long long> h(static_cast<size_t>(n) * kScoreStride);
  const size_t first = static_cast<size_t>(slot) * p->hist_len * kScoreStride;
  HIP_TRY(hipMemcpy(h.data(), p->hist.as<long long>() + first, h.size() * sizeof(long long),
                    hipMemcpyDeviceToHost));
  const double inv_nrg = std::ldexp(1.0, -p->cfg.nrg_bits);
  for (int i = 0; i < n; i++) {
    if (energy != nullptr) {
      energy[i] = static_cast<double>(h[static_cast<size_t>(i) * kScoreStride + kPotentialSlot]) *
                  inv_nrg;
    }
    if (rows != nullptr) {
      std::memcpy(rows + static_cast<size_t>(i) * kScoreStride,
                  h.data() + static_cast<size_t>(i) * kScoreStride,
                  kScoreStride * sizeof(long long));
    }
  }
  *count = n;
  return STORMM_OK;
}

extern "C" int stormm_pool_coordinates(const stormm_pool *pool, int32_t slot, double *xyz,
                                       int64_t *fixed_xyz) {
  if (pool == nullptr || slot < 0 || slot >= pool->cap_slots || !pool->slot_live[slot]) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "slot is not resident");
  }
  const int n = pool->slot_natom[slot];
  const int s0 = pool->slot_start[slot];
  std::vector<long long> x(n), y(n), z(n);
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(x.data(), pool->xcrd.as<long long>() + s0, n * sizeof(long long),
                    hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(y.data(), pool->ycrd.as<long long>() + s0, n * sizeof(long long),
                    hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(z.data(), pool->zcrd.as<long long>() + s0, n * sizeof(long long),
                    hipMemcpyDeviceToHost));
  const double inv = std::ldexp(1.0, -pool->cfg.gpos_bits);
  for (int i = 0; i < n; i++) {
    if (xyz) {
      xyz[3 * i] = static_cast<double>(x[i]) * inv;
      xyz[3 * i + 1] = static_cast<double>(y[i]) * inv;
      xyz[3 * i + 2] = static_cast<double>(z[i]) * inv;
    }
    if (fixed_xyz) {
      fixed_xyz[3 * i] = x[i];
      fixed_xyz[3 * i + 1] = y[i];
      fixed_xyz[3 * i + 2] = z[i];
    }
  }
  return STORMM_OK;
}
