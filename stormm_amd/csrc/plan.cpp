// Host planning of batched evaluations (see plan.h).
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

namespace stormm_amd {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Per-type sigma and epsilon when the table follows Lorentz-Berthelot combining.
static bool lorentz_berthelot(int t, const double *A, const double *B, const double *A14,
                              const double *B14, std::vector<double> &sig,
                              std::vector<double> &eps) {
  sig.assign(t, 0.0);
  eps.assign(t, 0.0);
  for (int i = 0; i < t * t; i++) {
    if (A[i] != A14[i] || B[i] != B14[i]) {
      return false;
    }
  }
  for (int i = 0; i < t; i++) {
    const double a = A[i * t + i], b = B[i * t + i];
    if (a > 0.0 && b > 0.0) {
      sig[i] = std::pow(a / b, 1.0 / 6.0);
      eps[i] = (b * b) / (4.0 * a);
    }
    else if (a != 0.0 || b != 0.0) {
      return false;
    }
  }
  for (int i = 0; i < t; i++) {
    for (int j = 0; j < t; j++) {
      const double s = 0.5 * (sig[i] + sig[j]);
      const double e = std::sqrt(eps[i] * eps[j]);
      const double s2 = s * s;
      const double s6 = s2 * s2 * s2;
      const double a_lb = 4.0 * e * s6 * s6;
      const double b_lb = 4.0 * e * s6;
      const double a = A[i * t + j], b = B[i * t + j];
      if (std::fabs(a_lb - a) > 1.0e-10 * std::max(std::fabs(a), 1.0e-300) ||
          std::fabs(b_lb - b) > 1.0e-10 * std::max(std::fabs(b), 1.0e-300)) {
        return false;
      }
    }
  }
  return true;
}


// FNV-1a over the parts of a system's topology that its work template depends on.
struct Fnv {
  unsigned long long h = 1469598103934665603ULL;
  void bytes(const void *p, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; i++) {
      h = (h ^ c[i]) * 1099511628211ULL;
    }
  }
  template <typename T> void span(const T *p, size_t n) {
    if (p != nullptr && n > 0) {
      bytes(p, n * sizeof(T));
    }
  }
};

static unsigned long long topology_hash(const stormm_topology_batch *b, const SystemTerms &t,
                                        int natom) {
  Fnv f;
  f.bytes(&natom, sizeof(natom));
  f.bytes(&t.lj_types, sizeof(t.lj_types));
  f.bytes(&t.combining, sizeof(t.combining));
  f.span(b->exclusion_pairs + 2 * t.excl_first, 2 * static_cast<size_t>(t.excl_count));
  f.span(b->bond_atoms + 2 * t.bond_first, 2 * static_cast<size_t>(t.bond_count));
  f.span(b->bond_k + t.bond_first, t.bond_count);
  f.span(b->bond_r0 + t.bond_first, t.bond_count);
  f.span(b->angle_atoms + 3 * t.angl_first, 3 * static_cast<size_t>(t.angl_count));
  f.span(b->angle_k + t.angl_first, t.angl_count);
  f.span(b->angle_theta + t.angl_first, t.angl_count);
  f.span(b->dihedral_atoms + 4 * t.dihe_first, 4 * static_cast<size_t>(t.dihe_count));
  f.span(b->dihe_amp + t.dihe_first, t.dihe_count);
  f.span(b->dihe_phase + t.dihe_first, t.dihe_count);
  f.span(b->dihe_freq + t.dihe_first, t.dihe_count);
  f.span(b->dihe_improper + t.dihe_first, t.dihe_count);
  f.span(b->dihe_scee + t.dihe_first, t.dihe_count);
  f.span(b->dihe_scnb + t.dihe_first, t.dihe_count);
  if (t.combining) {
    f.span(b->lj_sigma + t.sig_first, t.lj_types);
    f.span(b->lj_epsilon + t.sig_first, t.lj_types);
  }
  else {
    const size_t t2 = static_cast<size_t>(t.lj_types) * t.lj_types;
    f.span(b->lja + t.lj_first, t2);
    f.span(b->ljb + t.lj_first, t2);
    f.span(b->lja14 + t.lj_first, t2);
    f.span(b->ljb14 + t.lj_first, t2);
  }
  return f.h;
}

template <typename T>
static bool same_span(const T *base, int first_a, int first_b, size_t n) {
  if (n == 0 || base == nullptr) {
    return true;
  }
  return std::memcmp(base + first_a, base + first_b, n * sizeof(T)) == 0;
}

static bool same_topology(const stormm_topology_batch *b, const SystemTerms &x, int nx,
                          const SystemTerms &y, int ny) {
  if (nx != ny || x.bond_count != y.bond_count || x.angl_count != y.angl_count ||
      x.dihe_count != y.dihe_count || x.excl_count != y.excl_count ||
      x.lj_types != y.lj_types || x.combining != y.combining) {
    return false;
  }
  const size_t t2 = x.combining ? 0 : static_cast<size_t>(x.lj_types) * x.lj_types;
  const size_t t1 = x.combining ? static_cast<size_t>(x.lj_types) : 0;
  return same_span(b->lj_sigma, x.sig_first, y.sig_first, t1) &&
         same_span(b->lj_epsilon, x.sig_first, y.sig_first, t1) &&
         same_span(b->exclusion_pairs, 2 * x.excl_first, 2 * y.excl_first, 2 * x.excl_count) &&
         same_span(b->bond_atoms, 2 * x.bond_first, 2 * y.bond_first, 2 * x.bond_count) &&
         same_span(b->bond_k, x.bond_first, y.bond_first, x.bond_count) &&
         same_span(b->bond_r0, x.bond_first, y.bond_first, x.bond_count) &&
         same_span(b->angle_atoms, 3 * x.angl_first, 3 * y.angl_first, 3 * x.angl_count) &&
         same_span(b->angle_k, x.angl_first, y.angl_first, x.angl_count) &&
         same_span(b->angle_theta, x.angl_first, y.angl_first, x.angl_count) &&
         same_span(b->dihedral_atoms, 4 * x.dihe_first, 4 * y.dihe_first, 4 * x.dihe_count) &&
         same_span(b->dihe_amp, x.dihe_first, y.dihe_first, x.dihe_count) &&
         same_span(b->dihe_phase, x.dihe_first, y.dihe_first, x.dihe_count) &&
         same_span(b->dihe_freq, x.dihe_first, y.dihe_first, x.dihe_count) &&
         same_span(b->dihe_improper, x.dihe_first, y.dihe_first, x.dihe_count) &&
         same_span(b->dihe_scee, x.dihe_first, y.dihe_first, x.dihe_count) &&
         same_span(b->dihe_scnb, x.dihe_first, y.dihe_first, x.dihe_count) &&
         same_span(b->lja, x.lj_first, y.lj_first, t2) &&
         same_span(b->ljb, x.lj_first, y.lj_first, t2) &&
         same_span(b->lja14, x.lj_first, y.lj_first, t2) &&
         same_span(b->ljb14, x.lj_first, y.lj_first, t2);
}

// Builds the work template of system s (tile tasks with exclusion masks, valence terms and
// their parameters, the LJ table) and appends it to the plan.
static int add_template(const stormm_topology_batch *b, const SystemTerms &t, int n, int s,
                        int batch_id, TemplateStore &P) {
  TopoTemplate tp;
  tp.natom = n;
  tp.batch_id = batch_id;
  tp.source_system = s;
  tp.bond_count = t.bond_count;
  tp.angl_count = t.angl_count;
  tp.dihe_count = t.dihe_count;
  tp.combining = t.combining;
  // Lennard-Jones: per-type Lorentz-Berthelot parameters, or a table that may follow the rule
  // (then per-atom parameters replace it) or must be read as stored.
  const int nt_lj = t.lj_types;
  tp.lj_nt = nt_lj;
  tp.lj_off = 0;
  if (t.combining) {
    tp.lb = true;
    tp.sig.assign(b->lj_sigma + t.sig_first, b->lj_sigma + t.sig_first + nt_lj);
    tp.eps.assign(b->lj_epsilon + t.sig_first, b->lj_epsilon + t.sig_first + nt_lj);
  }
  else {
    const double *A = b->lja + t.lj_first, *B = b->ljb + t.lj_first;
    const double *A14 = b->lja14 + t.lj_first, *B14 = b->ljb14 + t.lj_first;
    tp.lb = lorentz_berthelot(nt_lj, A, B, A14, B14, tp.sig, tp.eps);
    if (!tp.lb) {
      tp.lj_off = static_cast<int>(P.ljab.size());
      for (int k = 0; k < nt_lj * nt_lj; k++) {
        P.ljab.push_back(make_double2(A[k], B[k]));
        P.ljab14.push_back(make_double2(A14[k], B14[k]));
      }
    }
  }
  // Exclusion matrix (bit per pair): the given list, or, when the batch carries none, every pair
  // within three bonds (1:2, 1:3 and 1:4 by shortest bond path, the reference's exclusion map
  // for systems without virtual sites, src/Topology/atomgraph_refinement.cpp).
  std::vector<unsigned char> ex(static_cast<size_t>(n) * n, 0);
  if (b->exclusion_counts != nullptr) {
    for (int e = 0; e < t.excl_count; e++) {
      const int i = b->exclusion_pairs[2 * (t.excl_first + e)];
      const int j = b->exclusion_pairs[2 * (t.excl_first + e) + 1];
      if (i < 0 || j < 0 || i >= n || j >= n) {
        return fail(STORMM_ERR_INVALID_ARGUMENT, "exclusion index out of range in system " +
                                                     std::to_string(s));
      }
      ex[static_cast<size_t>(i) * n + j] = 1;
      ex[static_cast<size_t>(j) * n + i] = 1;
    }
  }
  else {
    std::vector<std::vector<int>> adj(n);
    for (int e = 0; e < t.bond_count; e++) {
      const int k = t.bond_first + e;
      const int i = b->bond_atoms[2 * k], j = b->bond_atoms[2 * k + 1];
      if (i < 0 || j < 0 || i >= n || j >= n) {
        return fail(STORMM_ERR_INVALID_ARGUMENT, "bond index out of range");
      }
      adj[i].push_back(j);
      adj[j].push_back(i);
    }
    std::vector<int> dist(n, -1), frontier, next, seen;
    for (int src = 0; src < n; src++) {
      frontier.assign(1, src);
      seen.assign(1, src);
      dist[src] = 0;
      for (int depth = 1; depth <= 3 && !frontier.empty(); depth++) {
        next.clear();
        for (int u : frontier) {
          for (int v : adj[u]) {
            if (dist[v] < 0) {
              dist[v] = depth;
              next.push_back(v);
              seen.push_back(v);
              ex[static_cast<size_t>(src) * n + v] = 1;
            }
          }
        }
        frontier.swap(next);
      }
      for (int v : seen) {
        dist[v] = -1;
      }
    }
  }
  // Tile tasks: the clean off-diagonal tiles (16 x 16 atoms, no excluded pair) first, then the
  // other off-diagonal tiles, then the diagonal ones.
  const int nt = round_up(n, kTile) / kTile;
  tp.task_first = static_cast<int>(P.tasks.size());
  for (int pass = 0; pass < 3; pass++) {
    for (int ti = 0; ti < nt; ti++) {
      for (int tj = ti; tj < nt; tj++) {
        const bool diag = (ti == tj);
        if (diag != (pass == 2)) {
          continue;
        }
        const int ni = std::min(kTile, n - ti * kTile);
        const int nj = std::min(kTile, n - tj * kTile);
        unsigned m[kTile];
        bool any = false;
        for (int il = 0; il < kTile; il++) {
          m[il] = 0u;
          if (il >= ni) {
            continue;
          }
          for (int jl = 0; jl < nj; jl++) {
            if (ex[static_cast<size_t>(ti * kTile + il) * n + (tj * kTile + jl)]) {
              m[il] |= (1u << jl);
              any = true;
            }
          }
        }
        if (!diag && ((pass == 0) != (!any && ni == kTile && nj == kTile))) {
          continue;
        }
        Task tk;
        tk.x = (ti * kTile) | ((tj * kTile) << 16);
        tk.y = ((diag ? 1 : 0) << 8) | (ni << 12) | (nj << 20);
        tk.z = -1;
        tk.w = 0;
        if (any) {
          tk.z = static_cast<int>(P.masks.size());
          for (int il = 0; il < kTile; il++) {
            P.masks.push_back(m[il]);
          }
        }
        P.tasks.push_back(tk);
        if (diag) {
          tp.diag_count++;
        }
        else {
          tp.offd_count++;
          tp.clean_count += (pass == 0) ? 1 : 0;
        }
      }
    }
  }
  // Valence terms with system-relative atom indices.
  tp.term_first = static_cast<int>(P.vterms.size());
  for (int e = 0; e < t.bond_count; e++) {
    const int k = t.bond_first + e;
    const int i = b->bond_atoms[2 * k], j = b->bond_atoms[2 * k + 1];
    if (i < 0 || j < 0 || i >= n || j >= n) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "bond index out of range");
    }
    P.vterms.push_back(make_int4(kVBond, i | (j << 16), 0,
                                 static_cast<int>(P.bond_par.size())));
    P.bond_par.push_back(make_double2(b->bond_k[k], std::fabs(b->bond_r0[k])));
  }
  for (int e = 0; e < t.angl_count; e++) {
    const int k = t.angl_first + e;
    const int i = b->angle_atoms[3 * k], j = b->angle_atoms[3 * k + 1];
    const int kk = b->angle_atoms[3 * k + 2];
    if (i < 0 || j < 0 || kk < 0 || i >= n || j >= n || kk >= n) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "angle index out of range");
    }
    P.vterms.push_back(make_int4(kVAngle, i | (j << 16), kk,
                                 static_cast<int>(P.angl_par.size())));
    P.angl_par.push_back(make_double2(b->angle_k[k], b->angle_theta[k]));
  }
  for (int e = 0; e < t.dihe_count; e++) {
    const int k = t.dihe_first + e;
    const int *q4 = &b->dihedral_atoms[4 * k];
    for (int m = 0; m < 4; m++) {
      if (q4[m] < 0 || q4[m] >= n) {
        return fail(STORMM_ERR_INVALID_ARGUMENT, "dihedral index out of range");
      }
    }
    const double freq = b->dihe_freq[k];
    if (!(freq >= 0.0 && freq <= kMaxDihePeriod && freq == std::rint(freq))) {
      return fail(STORMM_ERR_INVALID_ARGUMENT,
                  "dihedral periodicity must be an integer between 0 and 32");
    }
    P.vterms.push_back(make_int4(kVDihedral, q4[0] | (q4[1] << 16), q4[2] | (q4[3] << 16),
                                 static_cast<int>(P.dihe_par.size())));
    DiheParam dp;
    dp.amp = b->dihe_amp[k];
    dp.cos_phase = std::cos(b->dihe_phase[k]);
    dp.sin_phase = std::sin(b->dihe_phase[k]);
    dp.period = static_cast<int>(freq);
    dp.improper = b->dihe_improper[k] ? 1 : 0;
    P.dihe_par.push_back(dp);
    const double scee = b->dihe_scee[k], scnb = b->dihe_scnb[k];
    P.dihe14.push_back(make_double2(scee > 0.0 ? 1.0 / scee : 0.0,
                                    (scee > 0.0 && scnb > 0.0) ? 1.0 / scnb : 0.0));
  }
  tp.term_count = static_cast<int>(P.vterms.size()) - tp.term_first;
  tp.bond_count = t.bond_count;
  tp.angl_count = t.angl_count;
  tp.dihe_count = t.dihe_count;
  // Cost model: rotation steps of the three pair passes plus valence work.  A diagonal tile runs
  // 8 steps but counts 9: its wavefronts run the masked steps and take half as many pairs per
  // task, so set-up and sums weigh more (measured: the units pack better for it).
  tp.cost = (16.0 * tp.offd_count) + (9.0 * tp.diag_count) +
            0.25 * (t.bond_count + t.angl_count + 2 * t.dihe_count) + 0.1 * n;
  P.templates.push_back(std::move(tp));
  return STORMM_OK;
}


// Longest-processing-time assignment of systems onto `groups` x `per_group` work units, within
// the LDS atom capacity and the per-unit system limit; stored-table and combining-rule systems
// go to separate units.  Systems go, largest first, to the least loaded group and there to its
// least loaded unit that can hold them, so that the groups' totals come out even.  Units are
// emitted member by member: unit g + r * G is member r of group g, so the blocks that the
// hardware makes co-resident on one compute unit (one from each pass over the device) hold one
// group, and the compute units' loads are even rather than the units'.  Returns false (nothing
// emitted) when the capacities need more groups than there are compute units.
static bool pack_grouped(const std::vector<double> &cost, const std::vector<int> &natom,
                         const std::vector<char> &table_mode, int cap, int groups, int per_group,
                         std::vector<std::vector<int>> &bins_out, std::vector<int> &mode_out) {
  const int ns = static_cast<int>(cost.size());
  double total_cost = 0.0;
  for (int s = 0; s < ns; s++) {
    total_cost += cost[s];
  }
  struct Bin {
    double cost = 0.0;
    int atoms = 0;
    std::vector<int> systems;
  };
  struct Group {
    double cost = 0.0;
    int mode = 0;
    std::vector<Bin> bins;
  };
  std::vector<Group> all;
  for (int mode = 0; mode < 2; mode++) {
    std::vector<int> order;
    double mode_cost = 0.0;
    for (int s = 0; s < ns; s++) {
      if (table_mode[s] == mode) {
        order.push_back(s);
        mode_cost += cost[s];
      }
    }
    if (order.empty()) {
      continue;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
    const int target = std::max(1, static_cast<int>(std::lround(groups * mode_cost /
                                                                std::max(total_cost, 1.0))));
    std::vector<Group> grp;
    for (int s : order) {
      const int need = round_up(natom[s], kTile);
      int best_g = -1, best_b = -1;
      for (size_t g = 0; g < grp.size(); g++) {
        if (best_g >= 0 && grp[g].cost >= grp[best_g].cost) {
          continue;
        }
        int b_fit = -1;
        for (int b = 0; b < per_group; b++) {
          const Bin &bn = grp[g].bins[b];
          if (bn.atoms + need <= cap && static_cast<int>(bn.systems.size()) < kMaxWuSystems &&
              (b_fit < 0 || bn.cost < grp[g].bins[b_fit].cost)) {
            b_fit = b;
          }
        }
        if (b_fit >= 0) {
          best_g = static_cast<int>(g);
          best_b = b_fit;
        }
      }
      if (best_g < 0 || (static_cast<int>(grp.size()) < target && grp[best_g].cost > 0.0)) {
        Group ng;
        ng.mode = mode;
        ng.bins.resize(per_group);
        grp.push_back(std::move(ng));
        best_g = static_cast<int>(grp.size()) - 1;
        best_b = 0;
      }
      Bin &bn = grp[best_g].bins[best_b];
      bn.systems.push_back(s);
      bn.atoms += need;
      bn.cost += cost[s];
      grp[best_g].cost += cost[s];
    }
    for (Group &g : grp) {
      std::stable_sort(g.bins.begin(), g.bins.end(), [](const Bin &a, const Bin &c) {
        return a.cost > c.cost;
      });
      all.push_back(std::move(g));
    }
  }
  if (static_cast<int>(all.size()) > groups) {
    return false;
  }
  std::stable_sort(all.begin(), all.end(), [](const Group &a, const Group &c) {
    return a.cost > c.cost;
  });
  bins_out.clear();
  mode_out.clear();
  for (int r = 0; r < per_group; r++) {
    for (Group &g : all) {
      if (!g.bins[r].systems.empty()) {
        bins_out.push_back(std::move(g.bins[r].systems));
        mode_out.push_back(g.mode);
      }
    }
  }
  return true;
}

// Plain longest-processing-time assignment onto about `target_total` units, emitted in order of
// decreasing cost (for batches whose units outnumber the blocks the device holds at once, which
// the hardware then schedules as blocks finish).
static void pack_lpt(const std::vector<double> &cost, const std::vector<int> &natom,
                     const std::vector<char> &table_mode, int cap, int target_total,
                     std::vector<std::vector<int>> &bins_out, std::vector<int> &mode_out) {
  const int ns = static_cast<int>(cost.size());
  double total_cost = 0.0;
  for (int s = 0; s < ns; s++) {
    total_cost += cost[s];
  }
  struct Bin {
    double cost = 0.0;
    int atoms = 0;
    int mode = 0;
    std::vector<int> systems;
  };
  std::vector<Bin> bins;
  for (int mode = 0; mode < 2; mode++) {
    std::vector<int> order;
    double mode_cost = 0.0;
    for (int s = 0; s < ns; s++) {
      if (table_mode[s] == mode) {
        order.push_back(s);
        mode_cost += cost[s];
      }
    }
    if (order.empty()) {
      continue;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
    const int target = std::max(1, static_cast<int>(std::lround(target_total * mode_cost /
                                                                std::max(total_cost, 1.0))));
    const size_t first_bin = bins.size();
    for (int s : order) {
      const int need = round_up(natom[s], kTile);
      int best = -1;
      for (size_t k = first_bin; k < bins.size(); k++) {
        if (bins[k].atoms + need <= cap &&
            static_cast<int>(bins[k].systems.size()) < kMaxWuSystems &&
            (best < 0 || bins[k].cost < bins[best].cost)) {
          best = static_cast<int>(k);
        }
      }
      const bool open_new = (best < 0) ||
                            (static_cast<int>(bins.size() - first_bin) < target &&
                             bins[best].cost > 0.0);
      if (open_new) {
        Bin nb;
        nb.mode = mode;
        bins.push_back(nb);
        best = static_cast<int>(bins.size()) - 1;
      }
      bins[best].systems.push_back(s);
      bins[best].atoms += need;
      bins[best].cost += cost[s];
    }
  }
  std::stable_sort(bins.begin(), bins.end(), [](const Bin &a, const Bin &c) {
    return a.cost > c.cost;
  });
  bins_out.clear();
  mode_out.clear();
  for (Bin &bn : bins) {
    bins_out.push_back(std::move(bn.systems));
    mode_out.push_back(bn.mode);
  }
}

void pack_systems(const std::vector<double> &cost, const std::vector<int> &natom,
                  const std::vector<char> &table_mode, int cap, int groups, int per_group,
                  bool grouped, std::vector<std::vector<int>> &bins_out,
                  std::vector<int> &mode_out) {
  groups = std::max(1, groups);
  per_group = std::max(1, per_group);
  if (!grouped ||
      !pack_grouped(cost, natom, table_mode, cap, groups, per_group, bins_out, mode_out)) {
    pack_lpt(cost, natom, table_mode, cap, groups * per_group, bins_out, mode_out);
  }
}

int solvent_model(const stormm_topology_batch *b, SolventModel &sm) {
  if (b->igb != 0 && b->igb != 1 && b->igb != 2 && b->igb != 5) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "unsupported implicit solvent model (igb must be "
                                             "0, 1, 2 or 5)");
  }
  sm.igb = b->igb;
  sm.dielectric = (b->dielectric > 0.0) ? b->dielectric : 78.5;
  sm.kappa = std::sqrt(0.10806 * std::max(b->salt_concentration, 0.0));
  sm.coulomb = (b->coulomb_constant > 0.0) ? b->coulomb_constant : 332.0522172900000;
  // OBC alpha / beta / gamma, src/Constants/generalized_born.h:56-66.
  sm.alpha = sm.beta = sm.gamma = 0.0;
  if (b->igb == 2) {
    sm.alpha = 0.8;
    sm.gamma = 2.909125;
  }
  else if (b->igb == 5) {
    sm.alpha = 1.0;
    sm.beta = 0.8;
    sm.gamma = 4.85;
  }
  return STORMM_OK;
}

int collect_terms(const stormm_topology_batch *b, std::vector<SystemTerms> &st) {
  const int ns = b->system_count;
  st.assign(ns, SystemTerms());
  int ab = 0, an = 0, ad = 0, ae = 0, at = 0, asg = 0, atoms = 0;
  for (int s = 0; s < ns; s++) {
    const int n = b->atom_counts[s];
    if (n <= 0 || n > kMaxWuAtoms) {
      return fail(STORMM_ERR_CAPACITY, "system " + std::to_string(s) + " has " +
                                           std::to_string(n) + " atoms (supported: 1-512)");
    }
    st[s].atom_first = atoms;
    atoms += n;
    st[s].bond_first = ab;
    st[s].bond_count = b->bond_counts ? b->bond_counts[s] : 0;
    ab += st[s].bond_count;
    st[s].angl_first = an;
    st[s].angl_count = b->angle_counts ? b->angle_counts[s] : 0;
    an += st[s].angl_count;
    st[s].dihe_first = ad;
    st[s].dihe_count = b->dihedral_counts ? b->dihedral_counts[s] : 0;
    ad += st[s].dihe_count;
    st[s].excl_first = ae;
    st[s].excl_count = b->exclusion_counts ? b->exclusion_counts[s] : 0;
    ae += st[s].excl_count;
    st[s].lj_types = b->lj_type_counts[s];
    if (st[s].lj_types <= 0) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "system " + std::to_string(s) +
                                                   " has no Lennard-Jones types");
    }
    st[s].combining = (b->lj_combining != nullptr) && (b->lj_combining[s] != 0);
    if (st[s].combining) {
      if (b->lj_sigma == nullptr || b->lj_epsilon == nullptr) {
        return fail(STORMM_ERR_INVALID_ARGUMENT, "combining-rule systems need lj_sigma and "
                                                 "lj_epsilon");
      }
      st[s].sig_first = asg;
      asg += st[s].lj_types;
    }
    else {
      if (b->lja == nullptr || b->ljb == nullptr || b->lja14 == nullptr || b->ljb14 == nullptr) {
        return fail(STORMM_ERR_INVALID_ARGUMENT, "system " + std::to_string(s) +
                                                     " has no Lennard-Jones tables");
      }
      st[s].lj_first = at;
      at += st[s].lj_types * st[s].lj_types;
    }
  }
  return STORMM_OK;
}

int find_or_add_template(const stormm_topology_batch *b, const std::vector<SystemTerms> &st,
                         int s, int batch_id, TemplateStore &ts, int *tid) {
  const int n = b->atom_counts[s];
  const unsigned long long h = topology_hash(b, st[s], n);
  std::vector<int> &bucket = ts.by_hash[h];
  for (int cand : bucket) {
    const TopoTemplate &tp = ts.templates[cand];
    if (tp.batch_id == batch_id) {
      const int src = tp.source_system;
      if (same_topology(b, st[src], b->atom_counts[src], st[s], n)) {
        *tid = cand;
        return STORMM_OK;
      }
    }
    else if (tp.natom == n && tp.bond_count == st[s].bond_count &&
             tp.angl_count == st[s].angl_count && tp.dihe_count == st[s].dihe_count &&
             tp.lj_nt == st[s].lj_types && tp.combining == st[s].combining) {
      *tid = cand;
      return STORMM_OK;
    }
  }
  const int rc = add_template(b, st[s], n, s, batch_id, ts);
  if (rc != STORMM_OK) {
    return rc;
  }
  *tid = static_cast<int>(ts.templates.size()) - 1;
  ts.templates.back().hash = h;
  bucket.push_back(*tid);
  return STORMM_OK;
}

TemplateStoreMark mark_templates(const TemplateStore &ts) {
  TemplateStoreMark m;
  m.templates = ts.templates.size();
  m.tasks = ts.tasks.size();
  m.masks = ts.masks.size();
  m.vterms = ts.vterms.size();
  m.bond_par = ts.bond_par.size();
  m.angl_par = ts.angl_par.size();
  m.dihe14 = ts.dihe14.size();
  m.ljab = ts.ljab.size();
  m.ljab14 = ts.ljab14.size();
  m.dihe_par = ts.dihe_par.size();
  return m;
}

void rollback_templates(TemplateStore &ts, const TemplateStoreMark &m) {
  for (size_t t = m.templates; t < ts.templates.size(); t++) {
    auto it = ts.by_hash.find(ts.templates[t].hash);
    if (it == ts.by_hash.end()) {
      continue;
    }
    std::vector<int> &bucket = it->second;
    bucket.erase(std::remove(bucket.begin(), bucket.end(), static_cast<int>(t)), bucket.end());
    if (bucket.empty()) {
      ts.by_hash.erase(it);
    }
  }
  ts.templates.resize(m.templates);
  ts.tasks.resize(m.tasks);
  ts.masks.resize(m.masks);
  ts.vterms.resize(m.vterms);
  ts.bond_par.resize(m.bond_par);
  ts.angl_par.resize(m.angl_par);
  ts.dihe14.resize(m.dihe14);
  ts.ljab.resize(m.ljab);
  ts.ljab14.resize(m.ljab14);
  ts.dihe_par.resize(m.dihe_par);
}

size_t template_store_bytes(const TemplateStore &ts) {
  return ts.tasks.size() * sizeof(Task) + ts.masks.size() * sizeof(unsigned) +
         ts.vterms.size() * sizeof(int4) +
         (ts.bond_par.size() + ts.angl_par.size() + ts.dihe14.size() + ts.ljab.size() +
          ts.ljab14.size()) * sizeof(double2) +
         ts.dihe_par.size() * sizeof(DiheParam);
}

void compact_templates(TemplateStore &ts, const std::vector<char> &keep, std::vector<int> &remap) {
  TemplateStore out;
  remap.assign(ts.templates.size(), -1);
  for (size_t t = 0; t < ts.templates.size(); t++) {
    if (!keep[t]) {
      continue;
    }
    TopoTemplate tp = ts.templates[t];
    // Tile tasks and the exclusion masks of those that carry one (kTile words each).
    const int ntask = tp.offd_count + tp.diag_count;
    const int task_first = static_cast<int>(out.tasks.size());
    for (int i = 0; i < ntask; i++) {
      Task tk = ts.tasks[tp.task_first + i];
      if (tk.z >= 0) {
        const int mask_first = static_cast<int>(out.masks.size());
        out.masks.insert(out.masks.end(), ts.masks.begin() + tk.z,
                         ts.masks.begin() + tk.z + kTile);
        tk.z = mask_first;
      }
      out.tasks.push_back(tk);
    }
    tp.task_first = task_first;
    // Valence terms with their parameters (a dihedral's 1:4 factors share its index).
    const int term_first = static_cast<int>(out.vterms.size());
    for (int i = 0; i < tp.term_count; i++) {
      int4 v = ts.vterms[tp.term_first + i];
      if (v.x == kVBond) {
        out.bond_par.push_back(ts.bond_par[v.w]);
        v.w = static_cast<int>(out.bond_par.size()) - 1;
      }
      else if (v.x == kVAngle) {
        out.angl_par.push_back(ts.angl_par[v.w]);
        v.w = static_cast<int>(out.angl_par.size()) - 1;
      }
      else {
        out.dihe_par.push_back(ts.dihe_par[v.w]);
        out.dihe14.push_back(ts.dihe14[v.w]);
        v.w = static_cast<int>(out.dihe_par.size()) - 1;
      }
      out.vterms.push_back(v);
    }
    tp.term_first = term_first;
    if (!tp.lb) {
      const size_t nt2 = static_cast<size_t>(tp.lj_nt) * tp.lj_nt;
      const int lj_off = static_cast<int>(out.ljab.size());
      out.ljab.insert(out.ljab.end(), ts.ljab.begin() + tp.lj_off,
                      ts.ljab.begin() + tp.lj_off + nt2);
      out.ljab14.insert(out.ljab14.end(), ts.ljab14.begin() + tp.lj_off,
                        ts.ljab14.begin() + tp.lj_off + nt2);
      tp.lj_off = lj_off;
    }
    remap[t] = static_cast<int>(out.templates.size());
    out.by_hash[tp.hash].push_back(remap[t]);
    out.templates.push_back(std::move(tp));
  }
  ts = std::move(out);
}

int fill_atom_parameters(const stormm_topology_batch *b, const SystemTerms &t, int natom,
                         const TopoTemplate &tp, const SolventModel &sm, int start,
                         AtomParams &ap, int s) {
  const double sqrt_coul = std::sqrt(sm.coulomb);
  for (int i = 0; i < natom; i++) {
    const int g = start + i;
    const int src = t.atom_first + i;
    const int ty = b->lj_type[src];
    if (ty < 0 || ty >= t.lj_types) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "LJ type index out of range in system " +
                                                   std::to_string(s));
    }
    ap.q[g] = b->charge[src] * sqrt_coul;
    ap.lj_type[g] = ty;
    if (tp.lb) {
      ap.ljp1[g] = tp.sig[ty];
      ap.ljp2[g] = 2.0 * std::sqrt(tp.eps[ty]);
    }
    const double pr = (b->pb_radii != nullptr) ? b->pb_radii[src] : 0.0;
    const double sc = (b->gb_screen != nullptr) ? b->gb_screen[src] : 0.0;
    ap.pbr[g] = pr;
    ap.rho[g] = pr - 0.09;
    ap.scr[g] = sc * (pr - 0.09);
    ap.alpha[g] = sm.alpha;
    ap.beta[g] = sm.beta;
    ap.gamma[g] = sm.gamma;
    if (sm.igb != 0 && ap.rho[g] <= 0.0) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "Born radius must exceed 0.09 A (system " +
                                                   std::to_string(s) + ")");
    }
  }
  return STORMM_OK;
}

int append_work_unit(const std::vector<int> &systems, int lj_table,
                     const std::vector<int> &sys_index, const std::vector<int> &atom_starts,
                     const std::vector<int> &atom_counts, const std::vector<int> &sys_template,
                     const TemplateStore &ts, std::vector<WuDesc> &wu,
                     std::vector<WuSys> &wu_sys, const std::vector<int> *rst_first,
                     const std::vector<int> *rst_count) {
  WuDesc d;
  d.sys_first = static_cast<int>(wu_sys.size());
  d.sys_count = static_cast<int>(systems.size());
  d.offd_count = 0;
  d.task_count = 0;
  d.term_count = 0;
  d.lj_table = lj_table;
  d.rst_count = 0;
  int lds = 0;
  for (int s : systems) {
    const TopoTemplate &tp = ts.templates[sys_template[s]];
    WuSys ws;
    ws.sys_index = sys_index[s];
    ws.lds_base = lds;
    ws.global_start = atom_starts[s];
    ws.natom = atom_counts[s];
    ws.task_first = tp.task_first;
    ws.offd_count = tp.offd_count;
    ws.clean_count = tp.clean_count;
    ws.diag_count = tp.diag_count;
    ws.term_first = tp.term_first;
    ws.term_count = tp.term_count;
    ws.bond_count = tp.bond_count;
    ws.angl_count = tp.angl_count;
    ws.lj_off = tp.lj_off;
    ws.lj_nt = tp.lj_nt;
    ws.rst_first = (rst_first != nullptr) ? (*rst_first)[s] : 0;
    ws.rst_count = (rst_count != nullptr) ? (*rst_count)[s] : 0;
    wu_sys.push_back(ws);
    d.offd_count += tp.offd_count;
    d.task_count += tp.offd_count + tp.diag_count;
    d.term_count += tp.term_count;
    d.rst_count += ws.rst_count;
    lds += round_up(atom_counts[s], kTile);
  }
  d.lds_atoms = lds;
  wu.push_back(d);
  return lds;
}

int collect_restraints(const stormm_topology_batch *b, std::vector<stormm_restraint> &rst,
                       std::vector<int> &first, std::vector<int> &count) {
  const int ns = b->system_count;
  rst.clear();
  first.assign(ns, 0);
  count.assign(ns, 0);
  if (b->restraint_counts == nullptr) {
    return STORMM_OK;
  }
  if (b->restraints == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "restraint counts given without restraints");
  }
  static const int natoms_of_kind[5] = {0, 1, 2, 3, 4};
  size_t pos = 0;
  for (int s = 0; s < ns; s++) {
    const int n = b->restraint_counts[s];
    if (n < 0) {
      return fail(STORMM_ERR_INVALID_ARGUMENT, "negative restraint count");
    }
    first[s] = static_cast<int>(rst.size());
    count[s] = n;
    for (int r = 0; r < n; r++) {
      const stormm_restraint &t = b->restraints[pos++];
      const std::string where = "restraint " + std::to_string(r) + " of system " +
                                std::to_string(s);
      if (t.kind < STORMM_RESTRAINT_POSITION || t.kind > STORMM_RESTRAINT_DIHEDRAL) {
        return fail(STORMM_ERR_INVALID_ARGUMENT, where + " has an unknown kind");
      }
      for (int a = 0; a < natoms_of_kind[t.kind]; a++) {
        if (t.atoms[a] < 0 || t.atoms[a] >= b->atom_counts[s]) {
          return fail(STORMM_ERR_INVALID_ARGUMENT, where + " names an atom outside the system");
        }
      }
      // src/Restraints/bounded_restraint.cpp:667-695 (the records arrive with the adjustments
      // of that routine already made).
      for (const double *r4 : {t.init_r, t.final_r}) {
        if (!(r4[0] <= r4[1] && r4[1] <= r4[2] && r4[2] <= r4[3])) {
          return fail(STORMM_ERR_INVALID_ARGUMENT, where + " needs r1 <= r2 <= r3 <= r4");
        }
        if (t.kind == STORMM_RESTRAINT_ANGLE && (r4[1] < 0.0 || r4[2] >= kPi)) {
          return fail(STORMM_ERR_INVALID_ARGUMENT, where + " needs r2 and r3 in [0, pi)");
        }
        if (t.kind == STORMM_RESTRAINT_DIHEDRAL && r4[2] - r4[1] >= 2.0 * kPi) {
          return fail(STORMM_ERR_INVALID_ARGUMENT, where + " needs r3 - r2 < 2 pi");
        }
      }
      if (t.final_step < t.init_step) {
        return fail(STORMM_ERR_INVALID_ARGUMENT, where + " ends before it starts");
      }
      rst.push_back(t);
    }
  }
  return STORMM_OK;
}

int build_plan(const stormm_topology_batch *b, const stormm_launch_shape *shape_in,
               SynthesisPlan &P) {
  if (b == nullptr || b->system_count <= 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "topology batch is empty");
  }
  int rc = solvent_model(b, P.sm);
  if (rc != STORMM_OK) {
    return rc;
  }
  stormm_launch_shape shape = {0, 0, 0};
  if (shape_in != nullptr) {
    shape = *shape_in;
  }
  const int cap = (shape.wu_atom_capacity > 0) ? shape.wu_atom_capacity : kMaxWuAtoms;
  if (cap < kTile || cap > kMaxWuAtoms || (cap % kTile) != 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "work-unit capacity must be a multiple of 16 in "
                                             "[16, 512]");
  }
  const int wu_per_cu = (shape.target_wu_per_cu > 0) ? shape.target_wu_per_cu : 2;
  P.block = (shape.block_threads > 0) ? shape.block_threads : kEvalBlock;
  if (P.block != 256 && P.block != 512 && P.block != 1024) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "block_threads must be 256, 512 or 1024");
  }
  const int ns = b->system_count;
  P.nsys = ns;
  std::vector<SystemTerms> st;
  rc = collect_terms(b, st);
  if (rc != STORMM_OK) {
    return rc;
  }
  rc = collect_restraints(b, P.rst, P.rst_first, P.rst_count);
  if (rc != STORMM_OK) {
    return rc;
  }

  // Global layout: each system starts on a 32-atom boundary.
  P.atom_starts.resize(ns);
  P.atom_counts.resize(ns);
  int run = 0;
  for (int s = 0; s < ns; s++) {
    const int n = b->atom_counts[s];
    if (round_up(n, kTile) > cap) {
      return fail(STORMM_ERR_CAPACITY, "system " + std::to_string(s) + " does not fit the "
                                           "work-unit capacity of " + std::to_string(cap));
    }
    P.atom_starts[s] = run;
    P.atom_counts[s] = n;
    run += round_up(n, kGlobalPad);
    P.pairs += static_cast<long long>(n) * (n - 1) / 2;
  }
  P.padded_atoms = run;
  P.ap.assign(run);

  // Templates (shared by systems of identical topology) and per-atom parameters.
  P.sys_template.assign(ns, -1);
  P.lj_table_mode.assign(ns, 0);
  for (int s = 0; s < ns; s++) {
    int tid = -1;
    rc = find_or_add_template(b, st, s, 0, P.ts, &tid);
    if (rc != STORMM_OK) {
      return rc;
    }
    P.sys_template[s] = tid;
    const TopoTemplate &tp = P.ts.templates[tid];
    P.lj_table_mode[s] = tp.lb ? 0 : 1;
    P.lj_table_systems += tp.lb ? 0 : 1;
    rc = fill_atom_parameters(b, st[s], b->atom_counts[s], tp, P.sm, P.atom_starts[s], P.ap, s);
    if (rc != STORMM_OK) {
      return rc;
    }
    P.task_count += tp.offd_count + tp.diag_count;
    P.vterm_count += tp.term_count + P.rst_count[s];
  }

  // Work units.
  const int cus = cu_count_or_default();
  std::vector<double> cost(ns);
  std::vector<int> sys_index(ns);
  for (int s = 0; s < ns; s++) {
    cost[s] = P.ts.templates[P.sys_template[s]].cost + P.rst_count[s];
    sys_index[s] = s;
  }
  std::vector<std::vector<int>> bins;
  std::vector<int> bin_mode;
  pack_systems(cost, P.atom_counts, P.lj_table_mode, cap, cus, wu_per_cu, true, bins, bin_mode);
  P.maxa = 0;
  for (size_t u = 0; u < bins.size(); u++) {
    P.maxa = std::max(P.maxa, append_work_unit(bins[u], bin_mode[u], sys_index, P.atom_starts,
                                               P.atom_counts, P.sys_template, P.ts, P.wu,
                                               P.wu_sys, &P.rst_first, &P.rst_count));
    if (P.wu.back().task_count > kMaxWuTasks) {
      return fail(STORMM_ERR_CAPACITY, "a work unit holds more than " +
                                           std::to_string(kMaxWuTasks) + " tile tasks");
    }
  }
  return STORMM_OK;
}

}  // namespace stormm_amd
