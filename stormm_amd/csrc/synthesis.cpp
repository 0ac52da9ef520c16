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
#include "internal.h"
#include "kernels.h"

using namespace stormm_amd;

// ------------------------------------------------------------------------------------------
// Error reporting
// ------------------------------------------------------------------------------------------
static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) {                                                                \
      return fail(STORMM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));      \
    }                                                                                      \
  } while (0)

extern "C" const char *stormm_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------------------------------
// Device buffers
// ------------------------------------------------------------------------------------------
struct DevMem {
  void *ptr = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem &) = delete;
  DevMem &operator=(const DevMem &) = delete;
  ~DevMem() {
    if (ptr != nullptr) {
      (void)hipFree(ptr);
    }
  }
  hipError_t alloc(size_t n) {
    bytes = n;
    if (n == 0) {
      return hipSuccess;
    }
    return hipMalloc(&ptr, n);
  }
  template <typename T> hipError_t upload(const std::vector<T> &v) {
    hipError_t e = alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) {
      return e;
    }
    return hipMemcpy(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// ------------------------------------------------------------------------------------------
// Host plan of a synthesis (no device memory)
// ------------------------------------------------------------------------------------------
// Offsets of one system's terms in the concatenated topology batch.
struct SystemTerms {
  int bond_first = 0, bond_count = 0;
  int angl_first = 0, angl_count = 0;
  int dihe_first = 0, dihe_count = 0;
  int excl_first = 0, excl_count = 0;
  int lj_types = 0;
  bool combining = false;  // per-type sigma / epsilon given (Lorentz-Berthelot), no tables
  int lj_first = 0;        // table block (T^2 entries) when !combining
  int sig_first = 0;       // per-type parameters (T entries) when combining
};

// Work template of one topology: its tile tasks (off-diagonal, then diagonal), exclusion masks,
// valence terms and parameters, with atom indices relative to the system, plus its LJ table.
struct TopoTemplate {
  int natom = 0;
  int source_system = 0;  // a system of the batch with this topology
  int task_first = 0, offd_count = 0, diag_count = 0;
  int term_first = 0, term_count = 0;
  int lj_off = 0, lj_nt = 0;
  bool lb = false;        // the LJ table follows Lorentz-Berthelot combining
  std::vector<double> sig, eps;
  double cost = 0.0;      // estimated evaluation cost (rotation steps of the pair passes)
};

struct SynthesisPlan {
  int nsys = 0;
  int padded_atoms = 0;
  int igb = 0;
  double dielectric = 78.5, kappa = 0.0, coulomb = 332.05221729;
  int maxa = 0;
  int block = kEvalBlock;
  std::vector<int> atom_starts, atom_counts;
  std::vector<double> q, ljp1, ljp2, rho, scr, pbr, alpha, beta, gamma;
  std::vector<int> lj_type;
  std::vector<double2> ljab, ljab14;
  std::vector<char> lj_table_mode;
  std::vector<TopoTemplate> templates;
  std::vector<int> sys_template;
  std::vector<WuDesc> wu;
  std::vector<WuSys> wu_sys;
  std::vector<Task> tasks;
  std::vector<unsigned> masks;
  std::vector<int4> vterms;
  std::vector<double2> bond_par, angl_par, dihe14;
  std::vector<DiheParam> dihe_par;
  long long pairs = 0;
  long long task_count = 0;
  long long vterm_count = 0;
  int lj_table_systems = 0;
};

static int round_up(int v, int m) { return ((v + m - 1) / m) * m; }

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

static int cu_count_or_default() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      prop.multiProcessorCount > 0) {
    return prop.multiProcessorCount;
  }
  return 256;
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
                        SynthesisPlan &P) {
  TopoTemplate tp;
  tp.natom = n;
  tp.source_system = s;
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
  // Tile tasks: off-diagonal tiles first, then the diagonal ones.
  const int nt = round_up(n, kTile) / kTile;
  tp.task_first = static_cast<int>(P.tasks.size());
  for (int pass = 0; pass < 2; pass++) {
    for (int ti = 0; ti < nt; ti++) {
      for (int tj = ti; tj < nt; tj++) {
        const bool diag = (ti == tj);
        if (diag != (pass == 1)) {
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
  // Cost model: rotation steps of the three pair passes plus valence work.
  tp.cost = (16.0 * tp.offd_count) + (9.0 * tp.diag_count) +
            0.25 * (t.bond_count + t.angl_count + 2 * t.dihe_count) + 0.1 * n;
  P.templates.push_back(std::move(tp));
  return STORMM_OK;
}

// Longest-processing-time assignment of systems onto about `target_total` work units, within
// the LDS atom capacity and the per-unit system limit; stored-table and combining-rule systems
// go to separate units.  Units come out in order of decreasing cost.
static void pack_systems(const std::vector<double> &cost, const std::vector<int> &natom,
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

// Appends one work unit holding `systems` to the plan; returns its LDS atom count.
static int append_work_unit(const std::vector<int> &systems, int lj_table, SynthesisPlan &P) {
  WuDesc d;
  d.sys_first = static_cast<int>(P.wu_sys.size());
  d.sys_count = static_cast<int>(systems.size());
  d.offd_count = 0;
  d.task_count = 0;
  d.term_count = 0;
  d.lj_table = lj_table;
  d.pad = 0;
  int lds = 0;
  for (int s : systems) {
    const TopoTemplate &tp = P.templates[P.sys_template[s]];
    WuSys ws;
    ws.sys_index = s;
    ws.lds_base = lds;
    ws.global_start = P.atom_starts[s];
    ws.natom = P.atom_counts[s];
    ws.task_first = tp.task_first;
    ws.offd_count = tp.offd_count;
    ws.diag_count = tp.diag_count;
    ws.term_first = tp.term_first;
    ws.term_count = tp.term_count;
    ws.lj_off = tp.lj_off;
    ws.lj_nt = tp.lj_nt;
    ws.pad = 0;
    P.wu_sys.push_back(ws);
    d.offd_count += tp.offd_count;
    d.task_count += tp.offd_count + tp.diag_count;
    d.term_count += tp.term_count;
    lds += round_up(P.atom_counts[s], kTile);
  }
  d.lds_atoms = lds;
  P.wu.push_back(d);
  return lds;
}

static int build_plan(const stormm_topology_batch *b, const stormm_launch_shape *shape_in,
                      SynthesisPlan &P) {
  if (b == nullptr || b->system_count <= 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "topology batch is empty");
  }
  if (b->igb != 0 && b->igb != 1 && b->igb != 2 && b->igb != 5) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "unsupported implicit solvent model (igb must be "
                                             "0, 1, 2 or 5)");
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
  const int ns = b->system_count;
  P.nsys = ns;
  P.igb = b->igb;
  P.dielectric = (b->dielectric > 0.0) ? b->dielectric : 78.5;
  P.kappa = std::sqrt(0.10806 * std::max(b->salt_concentration, 0.0));
  P.coulomb = (b->coulomb_constant > 0.0) ? b->coulomb_constant : 332.0522172900000;
  const double sqrt_coul = std::sqrt(P.coulomb);

  // Global layout: each system starts on a 32-atom boundary.
  P.atom_starts.resize(ns);
  P.atom_counts.resize(ns);
  int run = 0;
  for (int s = 0; s < ns; s++) {
    const int n = b->atom_counts[s];
    if (n <= 0 || n > kMaxWuAtoms) {
      return fail(STORMM_ERR_CAPACITY, "system " + std::to_string(s) + " has " +
                                           std::to_string(n) + " atoms (supported: 1-512)");
    }
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
  const int pa = run;
  P.q.assign(pa, 0.0);
  P.ljp1.assign(pa, 0.0);
  P.ljp2.assign(pa, 0.0);
  P.rho.assign(pa, 0.0);
  P.scr.assign(pa, 0.0);
  P.pbr.assign(pa, 0.0);
  P.alpha.assign(pa, 0.0);
  P.beta.assign(pa, 0.0);
  P.gamma.assign(pa, 0.0);
  P.lj_type.assign(pa, 0);
  double ga = 0.0, gb = 0.0, gg = 0.0;
  if (b->igb == 2) {
    ga = 0.8; gb = 0.0; gg = 2.909125;
  }
  else if (b->igb == 5) {
    ga = 1.0; gb = 0.8; gg = 4.85;
  }

  // Term offsets per system.
  std::vector<SystemTerms> st(ns);
  {
    int ab = 0, an = 0, ad = 0, ae = 0, at = 0, asg = 0;
    for (int s = 0; s < ns; s++) {
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
        if (b->lja == nullptr || b->ljb == nullptr || b->lja14 == nullptr ||
            b->ljb14 == nullptr) {
          return fail(STORMM_ERR_INVALID_ARGUMENT, "system " + std::to_string(s) +
                                                       " has no Lennard-Jones tables");
        }
        st[s].lj_first = at;
        at += st[s].lj_types * st[s].lj_types;
      }
    }
  }

  // Topology templates: systems with identical atom counts, exclusions, valence terms and LJ
  // tables share one set of tile tasks, exclusion masks, valence terms and parameters.
  P.sys_template.assign(ns, -1);
  std::unordered_map<unsigned long long, std::vector<int>> by_hash;
  for (int s = 0; s < ns; s++) {
    const unsigned long long h = topology_hash(b, st[s], b->atom_counts[s]);
    std::vector<int> &bucket = by_hash[h];
    int tid = -1;
    for (int cand : bucket) {
      const int src = P.templates[cand].source_system;
      if (same_topology(b, st[src], b->atom_counts[src], st[s], b->atom_counts[s])) {
        tid = cand;
        break;
      }
    }
    if (tid < 0) {
      const int rc = add_template(b, st[s], b->atom_counts[s], s, P);
      if (rc != STORMM_OK) {
        return rc;
      }
      tid = static_cast<int>(P.templates.size()) - 1;
      bucket.push_back(tid);
    }
    P.sys_template[s] = tid;
  }

  // Per-atom parameters.
  P.lj_table_mode.assign(ns, 0);
  int atom_cursor = 0;
  for (int s = 0; s < ns; s++) {
    const TopoTemplate &tp = P.templates[P.sys_template[s]];
    P.lj_table_mode[s] = tp.lb ? 0 : 1;
    P.lj_table_systems += tp.lb ? 0 : 1;
    const int t = st[s].lj_types;
    const int n = b->atom_counts[s];
    for (int i = 0; i < n; i++) {
      const int g = P.atom_starts[s] + i;
      const int src = atom_cursor + i;
      const int ty = b->lj_type[src];
      if (ty < 0 || ty >= t) {
        return fail(STORMM_ERR_INVALID_ARGUMENT, "LJ type index out of range in system " +
                                                     std::to_string(s));
      }
      P.q[g] = b->charge[src] * sqrt_coul;
      P.lj_type[g] = ty;
      if (tp.lb) {
        P.ljp1[g] = tp.sig[ty];
        P.ljp2[g] = 2.0 * std::sqrt(tp.eps[ty]);
      }
      const double pr = (b->pb_radii != nullptr) ? b->pb_radii[src] : 0.0;
      const double sc = (b->gb_screen != nullptr) ? b->gb_screen[src] : 0.0;
      P.pbr[g] = pr;
      P.rho[g] = pr - 0.09;
      P.scr[g] = sc * (pr - 0.09);
      P.alpha[g] = ga;
      P.beta[g] = gb;
      P.gamma[g] = gg;
      if (b->igb != 0 && P.rho[g] <= 0.0) {
        return fail(STORMM_ERR_INVALID_ARGUMENT, "Born radius must exceed 0.09 A (system " +
                                                     std::to_string(s) + ")");
      }
    }
    atom_cursor += n;
    P.task_count += tp.offd_count + tp.diag_count;
    P.vterm_count += tp.term_count;
  }

  // Work-unit packing: longest-processing-time assignment onto a target number of units,
  // respecting the LDS atom capacity and the per-unit system limit; systems that read stored
  // LJ tables are packed separately from combining-rule systems.
  const int cus = cu_count_or_default();
  std::vector<double> cost(ns);
  for (int s = 0; s < ns; s++) {
    cost[s] = P.templates[P.sys_template[s]].cost;
  }
  std::vector<std::vector<int>> bins;
  std::vector<int> bin_mode;
  pack_systems(cost, P.atom_counts, P.lj_table_mode, cap, std::max(1, cus * wu_per_cu), bins,
               bin_mode);
  P.maxa = 0;
  for (size_t u = 0; u < bins.size(); u++) {
    P.maxa = std::max(P.maxa, append_work_unit(bins[u], bin_mode[u], P));
  }
  P.block = (shape.block_threads > 0) ? shape.block_threads : kEvalBlock;
  if (P.block != 256 && P.block != 512 && P.block != 1024) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "block_threads must be 256, 512 or 1024");
  }
  return STORMM_OK;
}

// ------------------------------------------------------------------------------------------
// Objects
// ------------------------------------------------------------------------------------------
struct stormm_synthesis {
  SynthesisPlan plan;
  DevMem atom_starts, atom_counts, q, ljp1, ljp2, lj_type, ljab, ljab14, rho, scr,
      pbr, alpha, beta, gamma, wu, wu_sys, tasks, masks, vterms, bond_par, angl_par, dihe_par,
      dihe14, status, cycles_run, rms_force;
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
  HIP_TRY(ag->q.upload(P.q));
  HIP_TRY(ag->ljp1.upload(P.ljp1));
  HIP_TRY(ag->ljp2.upload(P.ljp2));
  HIP_TRY(ag->lj_type.upload(P.lj_type));
  HIP_TRY(ag->ljab.upload(P.ljab));
  HIP_TRY(ag->ljab14.upload(P.ljab14));
  HIP_TRY(ag->rho.upload(P.rho));
  HIP_TRY(ag->scr.upload(P.scr));
  HIP_TRY(ag->pbr.upload(P.pbr));
  HIP_TRY(ag->alpha.upload(P.alpha));
  HIP_TRY(ag->beta.upload(P.beta));
  HIP_TRY(ag->gamma.upload(P.gamma));
  HIP_TRY(ag->wu.upload(P.wu));
  HIP_TRY(ag->wu_sys.upload(P.wu_sys));
  HIP_TRY(ag->tasks.upload(P.tasks));
  HIP_TRY(ag->masks.upload(P.masks));
  HIP_TRY(ag->vterms.upload(P.vterms));
  HIP_TRY(ag->bond_par.upload(P.bond_par));
  HIP_TRY(ag->angl_par.upload(P.angl_par));
  HIP_TRY(ag->dihe_par.upload(P.dihe_par));
  HIP_TRY(ag->dihe14.upload(P.dihe14));
  std::vector<int> zero_status(1, 0);
  HIP_TRY(ag->status.upload(zero_status));
  std::vector<int> zi(P.nsys, 0);
  std::vector<double> zd(P.nsys, 0.0);
  HIP_TRY(ag->cycles_run.upload(zi));
  HIP_TRY(ag->rms_force.upload(zd));
  MMKit &k = ag->kit;
  std::memset(&k, 0, sizeof(k));
  k.nsys = P.nsys;
  k.nwu = static_cast<int>(P.wu.size());
  k.maxa = round_up(std::max(P.maxa, kTile), kTile);
  k.block = P.block;
  k.igb = P.igb;
  k.dielectric = P.dielectric;
  k.kappa = P.kappa;
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
  return k;
}

static int launch_terms(int32_t precision, const stormm_synthesis *ag, stormm_phasespace *ps,
                        stormm_scorecard *sc, int terms, int32_t eval_force, int32_t eval_energy,
                        void *stream) {
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
  const MMKit k = make_kit(ag, ps, sc);
  HIP_TRY(launch_mm_eval(precision == STORMM_PREC_DOUBLE ? 0 : 1, k, terms, eval_force ? 1 : 0,
                         eval_energy ? 1 : 0, static_cast<hipStream_t>(stream)));
  return STORMM_OK;
}

extern "C" int stormm_launch_nonbonded(int32_t precision, const stormm_synthesis *ag,
                                       stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                       stormm_scorecard *sc, int32_t eval_force,
                                       int32_t eval_energy, void *stream) {
  (void)ctrl;
  return launch_terms(precision, ag, ps, sc, kTermNonbonded, eval_force, eval_energy, stream);
}

extern "C" int stormm_launch_valence(int32_t precision, const stormm_synthesis *ag,
                                     stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                     stormm_scorecard *sc, int32_t eval_force,
                                     int32_t eval_energy, void *stream) {
  (void)ctrl;
  return launch_terms(precision, ag, ps, sc, kTermValence, eval_force, eval_energy, stream);
}

extern "C" int stormm_launch_mm_evaluation(int32_t precision, const stormm_synthesis *ag,
                                           stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                           stormm_scorecard *sc, int32_t eval_force,
                                           int32_t eval_energy, void *stream) {
  (void)ctrl;
  return launch_terms(precision, ag, ps, sc, kTermNonbonded | kTermValence, eval_force,
                      eval_energy, stream);
}

static CgKit make_cg_kit(const stormm_synthesis *ag, stormm_phasespace *ps, int precision) {
  CgKit k;
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
  stormm_mm_controls ctrl;
  ctrl.mc = *mc;
  ctrl.step = 0;
  int rc = stormm_phasespace_prime_conjugate_gradient(ps, stream);
  if (rc != STORMM_OK) {
    return rc;
  }
  const int prec = precision;
  for (int cyc = 0; cyc < mc->total_cycles; cyc++) {
    if ((rc = stormm_phasespace_initialize_forces(ps, stream)) != STORMM_OK ||
        (rc = stormm_scorecard_initialize(sc, stream)) != STORMM_OK ||
        (rc = stormm_launch_mm_evaluation(prec, ag, &ctrl, ps, sc, 1, 1, stream)) != STORMM_OK) {
      return rc;
    }
    ctrl.step += 1;
    if ((rc = stormm_launch_conjugate_gradient(prec, ag, ps, &ctrl, stream)) != STORMM_OK) {
      return rc;
    }
    for (int move = 0; move < 4; move++) {
      if (move > 0) {
        if ((rc = stormm_scorecard_initialize(sc, stream)) != STORMM_OK ||
            (rc = stormm_launch_mm_evaluation(prec, ag, &ctrl, ps, sc, 0, 1, stream)) !=
                STORMM_OK) {
          return rc;
        }
      }
      if ((rc = stormm_launch_line_advance(prec, ag, ps, sc, lm, move, stream)) != STORMM_OK) {
        return rc;
      }
    }
  }
  if ((rc = stormm_scorecard_initialize(sc, stream)) != STORMM_OK ||
      (rc = stormm_launch_mm_evaluation(prec, ag, &ctrl, ps, sc, 0, 1, stream)) != STORMM_OK) {
    return rc;
  }
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
  if (mincon->total_cycles < 0) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "negative cycle count");
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!persistent) {
    return minimize_stepwise(precision, ag, ps, mincon, sc, lm, st);
  }
  MinKit k;
  k.mm = make_kit(ag, ps, sc);
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
                                            int32_t precision, stormm_launch_shape *best,
                                            double *best_ms) {
  if (batch == nullptr || best == nullptr) {
    return fail(STORMM_ERR_INVALID_ARGUMENT, "null argument");
  }
  int maxn = 0;
  for (int s = 0; s < batch->system_count; s++) {
    maxn = std::max(maxn, batch->atom_counts[s]);
  }
  const int caps[4] = {128, 256, 384, 512};
  const int per_cu[3] = {1, 2, 4};
  const int blocks[3] = {256, 512, 1024};
  double best_t = 1.0e30;
  stormm_launch_shape best_shape = {512, 2, kEvalBlock};
  // Coordinates for timing: a loose cubic lattice per system (1.6 A spacing) keeps every pair
  // well separated; timing does not depend on the geometry beyond the far/near branch mix.
  std::vector<double> xyz;
  for (int s = 0; s < batch->system_count; s++) {
    for (int i = 0; i < batch->atom_counts[s]; i++) {
      xyz.push_back(1.6 * (i % 8));
      xyz.push_back(1.6 * ((i / 8) % 8));
      xyz.push_back(1.6 * (i / 64));
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
        rc = stormm_phasespace_upload_coordinates(ps, xyz.data());
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
