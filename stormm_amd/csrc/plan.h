// Host planning of batched evaluations: per-topology work templates (tile tasks with exclusion
// masks, valence terms and parameters, with atom indices relative to the system), the
// longest-processing-time packing of systems into work units, and the per-atom parameters.
// Used by the batch synthesis (synthesis.cpp) and the serving pool (pool.cpp).
#pragma once

#include <unordered_map>
#include <vector>

#include "../../include/stormm_amd.h"
#include "host_common.h"
#include "internal.h"

namespace stormm_amd {

// Offsets of one system's terms in a concatenated topology batch.
struct SystemTerms {
  int bond_first = 0, bond_count = 0;
  int angl_first = 0, angl_count = 0;
  int dihe_first = 0, dihe_count = 0;
  int excl_first = 0, excl_count = 0;
  int lj_types = 0;
  bool combining = false;  // per-type sigma / epsilon given (Lorentz-Berthelot), no tables
  int lj_first = 0;        // table block (T^2 entries) when !combining
  int sig_first = 0;       // per-type parameters (T entries) when combining
  int atom_first = 0;      // first atom of the system in the batch's per-atom arrays
};

// Work template of one topology.
struct TopoTemplate {
  int natom = 0;
  int batch_id = 0;       // the call that built it; -1 once that batch's arrays are gone
  int source_system = 0;  // a system of that batch with this topology
  int task_first = 0, offd_count = 0, clean_count = 0, diag_count = 0;
  int term_first = 0, term_count = 0;
  int bond_count = 0, angl_count = 0, dihe_count = 0;
  int lj_off = 0, lj_nt = 0;
  bool combining = false;
  bool lb = false;        // the LJ interactions follow Lorentz-Berthelot combining
  std::vector<double> sig, eps;
  double cost = 0.0;      // estimated evaluation cost (rotation steps of the pair passes)
  unsigned long long hash = 0;  // content hash, the key of TemplateStore::by_hash
};

// Templates and the arrays they index.
struct TemplateStore {
  std::vector<TopoTemplate> templates;
  std::vector<Task> tasks;
  std::vector<unsigned> masks;
  std::vector<int4> vterms;
  std::vector<double2> bond_par, angl_par, dihe14, ljab, ljab14;
  std::vector<DiheParam> dihe_par;
  std::unordered_map<unsigned long long, std::vector<int>> by_hash;
};

// Sizes of a TemplateStore's arrays, to undo the templates added after the mark.
struct TemplateStoreMark {
  size_t templates = 0, tasks = 0, masks = 0, vterms = 0, bond_par = 0, angl_par = 0,
         dihe14 = 0, ljab = 0, ljab14 = 0, dihe_par = 0;
};

TemplateStoreMark mark_templates(const TemplateStore &ts);

// Removes every template added after `m` (complete or partly built) and its array entries.
void rollback_templates(TemplateStore &ts, const TemplateStoreMark &m);

// Device bytes of a store's arrays.
size_t template_store_bytes(const TemplateStore &ts);

// Keeps only the templates with keep[t] != 0, packing their tasks, masks, valence terms and
// parameters to the front; remap[t] is template t's new index (-1 when dropped).
void compact_templates(TemplateStore &ts, const std::vector<char> &keep, std::vector<int> &remap);

// GB and electrostatic settings of a synthesis or pool.
struct SolventModel {
  int igb = 0;
  double dielectric = 78.5, kappa = 0.0, coulomb = 332.0522172900000;
  double alpha = 0.0, beta = 0.0, gamma = 0.0;
};

// Per-atom parameter arrays (one entry per padded global atom).
struct AtomParams {
  std::vector<double> q, ljp1, ljp2, rho, scr, pbr, alpha, beta, gamma;
  std::vector<int> lj_type;
  void assign(size_t n) {
    q.assign(n, 0.0);
    ljp1.assign(n, 0.0);
    ljp2.assign(n, 0.0);
    rho.assign(n, 0.0);
    scr.assign(n, 0.0);
    pbr.assign(n, 0.0);
    alpha.assign(n, 0.0);
    beta.assign(n, 0.0);
    gamma.assign(n, 0.0);
    lj_type.assign(n, 0);
  }
};

// Host plan of a batch synthesis (no device memory).
struct SynthesisPlan {
  int nsys = 0;
  int padded_atoms = 0;
  SolventModel sm;
  int maxa = 0;
  int block = kEvalBlock;
  std::vector<int> atom_starts, atom_counts;
  AtomParams ap;
  std::vector<char> lj_table_mode;
  TemplateStore ts;
  std::vector<int> sys_template;
  std::vector<WuDesc> wu;
  std::vector<WuSys> wu_sys;
  long long pairs = 0;
  long long task_count = 0;
  long long vterm_count = 0;
  int lj_table_systems = 0;
  std::vector<stormm_restraint> rst;       // all systems' restraints, system after system
  std::vector<int> rst_first, rst_count;   // per system
};

// Validated solvent settings of a topology batch.
int solvent_model(const stormm_topology_batch *b, SolventModel &sm);

// Term offsets of every system of a batch (validates counts and the LJ form).
int collect_terms(const stormm_topology_batch *b, std::vector<SystemTerms> &st);

// Template of system s of the batch, found among the store's templates or built and added.
// Templates built by earlier batches (batch_id < 0 or another id) match on the 64-bit content
// hash and the term counts; templates of the same batch are compared entry by entry.
int find_or_add_template(const stormm_topology_batch *b, const std::vector<SystemTerms> &st,
                         int s, int batch_id, TemplateStore &ts, int *tid);

// Per-atom parameters of system s written at global atoms [start, start + natom).
int fill_atom_parameters(const stormm_topology_batch *b, const SystemTerms &t, int natom,
                         const TopoTemplate &tp, const SolventModel &sm, int start,
                         AtomParams &ap, int s);

// Longest-processing-time assignment of systems onto about `groups` x `per_group` work units.
// With `grouped`, and when every unit fits on the device at once, the systems are balanced over
// `groups` groups (one per compute unit) emitted so that units g, g + G, g + 2G, ... form group
// g; otherwise the units are balanced individually and emitted by decreasing cost.
void pack_systems(const std::vector<double> &cost, const std::vector<int> &natom,
                  const std::vector<char> &table_mode, int cap, int groups, int per_group,
                  bool grouped, std::vector<std::vector<int>> &bins_out,
                  std::vector<int> &mode_out);

// Appends one work unit holding `systems` (indices into the per-system arrays); returns its
// LDS atom count.  `sys_index` gives each system's number in the kernels' per-system arrays;
// rst_first / rst_count (null: no restraints) each system's restraints.
int append_work_unit(const std::vector<int> &systems, int lj_table,
                     const std::vector<int> &sys_index, const std::vector<int> &atom_starts,
                     const std::vector<int> &atom_counts, const std::vector<int> &sys_template,
                     const TemplateStore &ts, std::vector<WuDesc> &wu,
                     std::vector<WuSys> &wu_sys, const std::vector<int> *rst_first = nullptr,
                     const std::vector<int> *rst_count = nullptr);

// Validated restraints of every system of a batch (atoms inside the system, bounds ordered).
int collect_restraints(const stormm_topology_batch *b, std::vector<stormm_restraint> &rst,
                       std::vector<int> &first, std::vector<int> &count);

// Whole plan of a batch synthesis.
int build_plan(const stormm_topology_batch *b, const stormm_launch_shape *shape_in,
               SynthesisPlan &P);

}  // namespace stormm_amd
